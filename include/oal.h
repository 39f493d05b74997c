/*
 * oal.h — C ABI of liboal, the MI355X (gfx950) D3Q27 cumulant lattice-Boltzmann hot path.
 *
 * Plain C: pointers, sizes and status codes only; no exceptions cross the boundary.  Every entry
 * point returns an int status: 0 (OAL_OK) on success, a negative OAL_ERR_* otherwise;
 * oal_strerror() names it and oal_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * The library owns all device memory of a plan.  Host arrays passed in use the reference's
 * BlockLevel conventions (src/geometry/blocks.jl:16-200): populations f[x,y,z,b,k] column-major
 * (x fastest, then y, z, block, direction), 1-based block coordinates map_x/y/z, and a
 * column-major n_blocks x 27 neighbour table of 1-based block indices with 0 = none
 * (src/geometry/domain_topology.jl:919-955).  Direction k follows the reference order
 * k = (cx+1) + 3 (cy+1) + 9 (cz+1) (src/collision/lattice.jl:10-59).
 *
 * Streams: every call that launches work takes a hipStream_t (passed as void*, NULL = the plan's
 * own stream) and returns without synchronising; oal_sync() waits for the plan's work.
 */
#ifndef OAL_H
#define OAL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OAL_ABI_VERSION 3

enum {
    OAL_OK = 0,
    OAL_ERR_ARG = -1,          /* invalid argument or descriptor */
    OAL_ERR_HIP = -2,          /* a HIP runtime call failed */
    OAL_ERR_NOMEM = -3,        /* device allocation failed */
    OAL_ERR_UNSUPPORTED = -4,  /* a feature outside this library's scope was requested */
    OAL_ERR_STATE = -5,        /* call out of order (e.g. no populations uploaded) */
    OAL_ERR_NONFINITE = -6     /* the stability guard found a bad cell */
};

/* population storage of a plan */
enum {
    OAL_F32 = 0,     /* FP32 populations, unshifted (the reference's storage) */
    OAL_F64 = 1,     /* FP64 populations and arithmetic (validation runs only) */
    OAL_F16S = 2,    /* FP16 storage of (f - w_k) * 2^14, FP32 arithmetic (the exact power-of-two
                        factor keeps the shifted values of low-Mach flows in FP16's normal range) */
    OAL_BF16S = 3    /* BF16 storage of f - w_k, FP32 arithmetic */
};

/* streaming pattern */
enum {
    OAL_STREAM_AB = 0,   /* pull from buffer A into buffer B, then swap (the reference's scheme) */
    OAL_STREAM_AA = 1    /* in-place AA pattern, one buffer, even/odd kernels */
};

/* rounding of the populations a step stores in the 16-bit forms (OAL_F16S, OAL_BF16S).  Stochastic
 * rounding draws its random bits from a hash of (cell, direction, time_step): results are
 * reproducible, and the same for every launch shape and streaming pattern.  Deterministic rounding
 * of populations that barely change from step to step repeats the same error every step, a steady
 * mass and momentum source that biases steady flows (sphere drag: 2 % in FP16, 25 % in BF16 at
 * Re = 200); stochastic rounding does not.  Uploads and initialisation round to nearest. */
enum {
    OAL_ROUND_STOCHASTIC = 0,
    OAL_ROUND_NEAREST = 1
};

/* treatment of obstacle cells */
enum {
    OAL_BB_REFERENCE = 0,  /* solid cells reflect their gathered populations (kernel_cumulant.jl:405-417) */
    OAL_BB_HALFWAY = 1     /* halfway bounce-back at the fluid cell, wall half a cell out */
};

/* interpolated bounce-back: the value of a q < 1/2 link whose cell behind (x - c_k) is not a fluid
 * cell of the level (BOUZIDI_FF_FALLBACK_MODE, src/io/config_loader.jl:2515-2527) */
enum {
    OAL_BFL_FALLBACK_HALFWAY = 0,      /* "halfway_bb": f_ff = f_k (the reference default) */
    OAL_BFL_FALLBACK_WALL_EQ = 1,      /* "wall_eq": f_ff = rho_f w_k */
    OAL_BFL_FALLBACK_BOUNCE_BACK = 2   /* "bounce_back": f_ff = f_opp(k) of the cell */
};

/* limiter_type of the reference kernel (kernel_cumulant.jl:725-735, 849-884) */
enum { OAL_LIMITER_NONE = 0, OAL_LIMITER_FACTORED = 1, OAL_LIMITER_POSITIVITY = 2 };

/* Collision parameters shared by every level (the cumulant arguments of stream_collide_cumulant!,
 * src/collision/kernel_cumulant.jl:42-76; defaults of src/io/config_loader.jl:82-88, 154-159). */
typedef struct {
    float omega_bulk;          /* CUMULANT_OMEGA_BULK (1.0) */
    float omega_3;             /* CUMULANT_OMEGA_3 (1.0) */
    float omega_4;             /* CUMULANT_OMEGA_4 (1.0) */
    float lambda;              /* CUMULANT_LAMBDA_PARAM (1/6) */
    int32_t adaptive_omega4;   /* CUMULANT_ADAPTIVE_OMEGA4 (1) */
    int32_t limiter;           /* OAL_LIMITER_* (factored) */
    int32_t compressible;      /* COMPRESSIBILITY_CORRECTION (0): O(u^4) equilibria at inlet/limiter */
    float c_wale;              /* WALE constant; 0 disables the subgrid model */
    float nu_sgs_background;   /* NU_SGS_BACKGROUND (0) */
    float inlet_turbulence;    /* INLET_TURBULENCE_INTENSITY; 0 = steady inlet */
    int32_t transverse_freestream;       /* TRANSVERSE_FREESTREAM_BOUNDARY (0): 1 = a population arriving
                                            through a y or z face is (1 - b) f_k + b feq(1, u_inlet)
                                            (kernel_cumulant.jl:323-339); excludes zwall_bc != 0 */
    float transverse_freestream_blend;   /* b, TRANSVERSE_FREESTREAM_BLEND (1), in [0, 1] */
} OalCollision;

/* One domain of a level: an extent in cells with its boundary conditions.  A level holds one
 * domain (uniform box, one octree level) or many (independent boxes advanced together). */
typedef struct {
    int32_t extent[3];         /* nx_global, ny_global, nz_global at this level */
    int32_t periodic[3];       /* nonzero: the axis wraps through the neighbour table */
    float tau;                 /* tau_molecular of this level */
    float u_inlet;             /* equilibrium inlet velocity along +x at x = 0 */
    float body_force[3];       /* uniform force density, lattice units of this level (Guo term) */
    float amr_sgs_scale;       /* _amr_sgs_scale (1 on a single level) */
    int32_t ymin_wall_bc;      /* 1: no-slip bounce-back at y_min (GROUND_ANALYSIS) */
    int32_t zwall_bc;          /* z faces (advanced.boundary.z_walls): 0 legacy mirror, 1 exact specular
                                  slip (the image cell's mirror_z population, kernel_cumulant.jl:272-321;
                                  one-level plans, AB pattern), 2 no-slip bounce-back */
    float force_scale;         /* factor on this domain's momentum-exchange sum (8^(1-l) on octree levels) */
    int32_t reserved;
} OalDomainDesc;

/* One refinement level (a BlockLevel).  Blocks of one domain must be contiguous in the list. */
typedef struct {
    int32_t n_blocks;
    const int32_t* map_x;            /* [n_blocks] 1-based block coordinates (active_coords_x) */
    const int32_t* map_y;
    const int32_t* map_z;
    const int32_t* neighbor_table;   /* [n_blocks * 27] column-major, 1-based, 0 = none */
    const int32_t* block_domain;     /* [n_blocks] 0-based domain index, NULL = all domain 0 */
    int32_t n_domains;
    const OalDomainDesc* domains;    /* [n_domains] */
    const uint8_t* obstacle;         /* [bs^3 * n_blocks] reference layout, nonzero = solid; NULL = none */
} OalLevelDesc;

/* A refined plan (n_levels > 1, levels coarsest first) is a block octree: level l + 1 has twice the
 * extent of level l, and a level-l block with children is refined into all 8 of them.  Requirements:
 * one domain per level (its tau in that level's lattice units; the acoustic rule is
 * tau_{l+1} - 1/2 = 2 (tau_l - 1/2), src/octree_q19/src/config.jl:328-335), extents multiples of 16,
 * the AB pattern, and a margin of one coarser block around every refined region.  The library derives
 * the neighbour tables of refined plans from the block coordinates, and adds a layer of ghost blocks
 * around every finer level.  With WALE, the ghost-cell fill and the restriction also write the velocity
 * of each cell they write, which the receiving level's gradient gather reads; each level's
 * amr_sgs_scale multiplies its eddy and background viscosity (the reference now uses 1 on every level,
 * src/collision/timestep.jl:1128). */
typedef struct {
    int32_t abi_version;       /* OAL_ABI_VERSION */
    int32_t storage;           /* OAL_F32 / OAL_F64 / OAL_F16S / OAL_BF16S */
    int32_t block_size;        /* cells per block edge: 16, or 8, the reference's BlockLevel
                                  (src/geometry/blocks.jl:14, BLOCK_SIZE = 8), for one-level plans:
                                  the library regroups each 2 x 2 x 2 of them into one of its 16^3
                                  blocks, which must then hold every octant that reaches into the
                                  domain extent (OAL_ERR_UNSUPPORTED otherwise); every per-cell host
                                  array (f, rho, u, q-map, sponge, LBM1 files) stays in the caller's
                                  8^3 layout.  The 16-bit forms' stochastic rounding draws its bits by
                                  the library's cell numbering. */
    int32_t stream_mode;       /* OAL_STREAM_AB or OAL_STREAM_AA */
    int32_t bb_mode;           /* OAL_BB_REFERENCE or OAL_BB_HALFWAY */
    int32_t n_levels;          /* 1 for a uniform grid */
    const OalLevelDesc* levels;
    OalCollision collision;
    int32_t device;            /* HIP device ordinal */
    int32_t rounding;          /* OAL_ROUND_STOCHASTIC (0) or OAL_ROUND_NEAREST: 16-bit forms only */
} OalPlanDesc;

typedef struct OalPlan OalPlan;

/* Replaces: nothing (ABI handshake). */
int oal_abi_version(void);
const char* oal_strerror(int status);
const char* oal_last_error(void);

/* Builds the device state of a solver run: the BlockLevel structs and their upload
 * (src/geometry/blocks.jl:250-507, Adapt at :203) and the lattice arrays of
 * build_lattice_arrays_gpu_d3q27 (src/collision/timestep.jl:2005-2023). */
int oal_plan_create(const OalPlanDesc* desc, OalPlan** plan);
int oal_plan_destroy(OalPlan* plan);

/* Number of cells (n_blocks * bs^3, the caller's blocks) of a level: the length of one direction of
 * the host arrays of f, rho and u. */
int oal_level_cells(const OalPlan* plan, int level, int64_t* n_cells);

/* Replaces: nothing (introspection).  How a level's blocks split over the step kernels: blocks on
 * a regular lexicographic grid with no boundary rule or solid link (n_regular, one arithmetic-address
 * launch; in the AB pattern also whole blocks on a domain face with no solid near), other blocks
 * with no boundary rule (n_interior), and the boundary blocks of both kinds plus the regular blocks
 * that overhang the extent (n_partial), launched on a second stream that overlaps the bulk launches.
 * The bulk launches that oal_time_steps times (ms_kernel) cover the n_regular and n_interior blocks. */
typedef struct OalLevelInfo {
    int64_t n_cells;         /* n_blocks * bs^3 (of the caller's blocks, 8^3 ones included; the
                                kernel counts below are of the library's 16^3 blocks) */
    int64_t n_fluid_cells;   /* inside the domain extent and not solid */
    int32_t n_blocks;
    int32_t n_regular;
    int32_t n_interior;
    int32_t n_boundary_regular;
    int32_t n_boundary;
    int32_t n_face_cells;    /* AB pattern: cells of regular blocks on a domain face; the bulk kernel
                                applies their rules itself or leaves them to the face-cell kernel
                                (OalTuning.face_kernel) */
    int32_t n_ghost_blocks;  /* refined plans: blocks the library adds around the level (filled from
                                the coarser level, not stepped; not part of n_blocks or n_cells) */
    int32_t n_covered_blocks;/* refined plans: blocks of n_blocks whose children exist (written by
                                restriction, not stepped) */
    int32_t n_partial;       /* AB pattern: regular blocks that overhang the extent, no solid near (the
                                pipelined kernel with the face rules at the extent's last cells, on the
                                second stream; in a domain whose block layers span 2 GiB or more they
                                are boundary blocks) */
    int32_t n_sponge;        /* AB pattern: regular blocks with sponge cells and no solid near (the
                                pipelined kernel's sponge form, on the second stream) */
    int32_t n_solid_regular; /* of n_boundary_regular: blocks whose only flag is a solid cell near, away
                                from the domain faces (FP32 arithmetic, halfway bounce-back: the
                                pipelined kernel's solid-link form on steps that do not collect the
                                force) */
} OalLevelInfo;
int oal_level_info(const OalPlan* plan, int level, OalLevelInfo* info);

/* Populations in/out in the reference layout f[x,y,z,b,k], unshifted; host_dtype OAL_F32 or
 * OAL_F64.  The library converts to its storage (and from it).  Replaces the copyto! of
 * BlockLevel.f / f_temp (src/geometry/blocks.jl:604-690, src/solver/solver_control.jl:830-851). */
int oal_upload_f(OalPlan* plan, int level, const void* host_f, int host_dtype);
int oal_download_f(OalPlan* plan, int level, void* host_f, int host_dtype);

/* Equilibrium initialisation at per-cell (rho, u): init_eq! (src/main.jl:546-572).
 * rho [n_cells] and u [3 * n_cells] in the reference layout ([x,y,z,b] and [x,y,z,b,3]), FP64;
 * NULL fields take rho0 / u0 everywhere. */
int oal_init_equilibrium(OalPlan* plan, int level, const double* rho, const double* u,
                         double rho0, const double u0[3]);

/* Per-step options of oal_level_step / oal_step */
enum {
    OAL_STEP_FORCE = 1         /* collect the momentum-exchange force of this step */
};

/* One step of one level: perform_timestep_v2! (src/collision/timestep.jl:946-2000) with the
 * cumulant operator of get_collision_kernel (src/collision/physics_kernels.jl:50-74).
 * time_step is the reference's time_step_seed (inlet noise) and keys the stochastic rounding of
 * the 16-bit forms: pass the step number, so that consecutive steps draw different bits. */
int oal_level_step(OalPlan* plan, int level, int32_t time_step, int flags, void* stream);

/* n coarse steps: execute_timestep_batch! (src/solver/solver_control.jl:796-824) over
 * recursive_step! (:55-422).  One level: n steps of level 0 with time_step first_time_step + i.
 * Refined plans (n_levels > 1): each coarse step steps level 0 once and, recursively, every finer
 * level twice per step of its parent (acoustic sub-cycling), with the octree's transfers
 * (src/octree_q19/src/coupling.jl:134-268): restriction into the covered cells after each fine
 * sub-step, ghost fill at the half and the full coarse time.  Level l's sub-step k of coarse step t
 * uses time_step 2^l t + k.  OAL_STEP_FORCE collects the force on the last sub-step of the last coarse
 * step of every level holding obstacle cells.  The first call after populations were set runs the
 * initial restriction and fill. */
int oal_step(OalPlan* plan, int n_steps, int32_t first_time_step, int flags, void* stream);

/* The momentum-exchange force of the last step collected with OAL_STEP_FORCE, FP64, per domain of
 * a level: F[3 * n_domains].  It sums, over every link from a fluid cell toward a solid cell, the
 * momentum the link carries in the streaming that follows the step: (f_i* + f_opp(i)(t+1)) c_i with
 * i toward the wall, shifted by w_i (src/octree_q19/src/stream.jl:248-256, reduce :382-389); the
 * reference collects the same sum while gathering the next step. */
int oal_force(OalPlan* plan, int level, double* force);

/* ---- Surface-stress forces on a triangle mesh (src/forces/surface.jl) ----------------------------
 * The reference's own drag path: each triangle probes density and velocity along its normal, turns
 * them into a pressure and a wall-law shear stress, and the stresses are integrated over the mesh.
 *
 * oal_set_surface uploads the mesh of a level (gpu_mesh of the reference): n_tri triangles as Float32
 * arrays centers[3 * n_tri] and normals[3 * n_tri] (x of every triangle, then y, then z; unit outward
 * normals) and areas[n_tri]; NULL or n_tri 0 removes it.  The level must have one domain.
 *
 * oal_surface_forces maps the stresses (map_surface_stresses! with map_stresses_kernel!, the default
 * Prandtl force method, surface.jl:636-1033 and 1333-1403, compute_stress_from_cell :180-269) and
 * integrates them (integrate_surface_forces!, :1265-1331 and 1452-1548), on `stream`, and waits.
 *   Probe field: the density and velocity of the level's current populations (the moments of
 *   oal_download_macro, rho clamped to >= 0.01 as the reference's kernel clamps it, then rounded to
 *   FP32); cells outside the extent, of absent blocks or solid are not sampled.  The mapping arithmetic
 *   is FP32 in the reference's order; each triangle's force and moment and their sums are FP64, summed
 *   in a fixed order (one result for one state, whatever the scheduling; the reference adds FP32
 *   atomics).  The positions are in the units of dx with the level's cell (i, j, k) centred at
 *   ((i + 1/2) dx, (j + 1/2) dx, (k + 1/2) dx) after mesh_offset is added to the centres.
 * oal_download_surface copies the per-triangle maps of the last oal_surface_forces (force_data's
 * pressure_map, shear_{x,y,z}_map, normal_sign_map and wall_dist_map; any pointer may be NULL):
 * pressure NaN where no fluid probe was found. */
typedef struct {
    double dx;                 /* cell size of the level in mesh units (level.dx) */
    double mesh_offset[3];     /* added to the triangle centres (params.mesh_offset) */
    double pressure_scale;     /* rho_physical * velocity_scale^2 (surface.jl:1343); 1 = lattice units */
    double stress_scale;       /* the same scale in the reference (surface.jl:1344) */
    double tau;                /* molecular relaxation time of the level (level.tau) */
    double c_wale;             /* C_SMAGO (src/io/config_loader.jl:82, 0.2): wall-law SGS estimate */
    double nu_sgs_bg;          /* NU_SGS_BACKGROUND (config_loader.jl:88, 0) */
    double cs2;                /* CS2_PHYSICS, 1/3 */
    double moment_center[3];   /* force_data.moment_center */
    double rho_ref, u_ref, area_ref, chord_ref;   /* coefficient references (force_data) */
    int32_t search_radius;     /* probes along the normal (main.jl:1296: max(8, ceil(0.12 / dx))) */
    int32_t recover_normal;    /* 1: probe the opposite side when the given one finds no fluid */
    int32_t symmetric;         /* 1: half model about y = 0 (force_data.symmetric): Fx, Fz, My doubled */
    int32_t reserved;
} OalSurfaceParams;

typedef struct {
    double F_pressure[3];      /* Fx_pressure, Fy_pressure, Fz_pressure */
    double F_viscous[3];       /* Fx_viscous, ... */
    double F[3];               /* their sum */
    double M[3];               /* moments about moment_center */
    double Cd, Cl, Cs;         /* Fx, Fz, Fy over q_inf area_ref, q_inf = rho_ref u_ref^2 / 2 (0 when that is <= 1e-10) */
    double Cm[3];              /* M over q_inf area_ref chord_ref */
    int64_t n_mapped;          /* triangles with a finite pressure and stress (those integrated) */
} OalSurfaceForces;

int oal_set_surface(OalPlan* plan, int level, int64_t n_tri, const float* centers, const float* normals,
                    const float* areas);
int oal_surface_forces(OalPlan* plan, int level, const OalSurfaceParams* params, OalSurfaceForces* out,
                       void* stream);
int oal_download_surface(OalPlan* plan, int level, float* pressure, float* shear, float* normal_sign,
                         float* wall_dist);

/* Interpolated (curved-wall) bounce-back on the solid links of a level: the linear Bouzidi rule
 * (BFL) of bouzidi_correction_kernel_fixed! (src/geometry/bouzidi_kernel.jl:150-380, launched by
 * apply_bouzidi_correction!, :382-432, from src/collision/timestep.jl:1752-1770) with its D3Q27
 * settings (no lattice-frame shift, no wall correction, no population limiter), and the Float16 q-map
 * of compute_bouzidi_qmap_sparse (src/geometry/bouzidi_setup.jl:67-163).
 *   q_map [n_cells * 27] Float16 bits in the reference layout q_map[x,y,z,b,k]: the distance from
 *         the cell centre to the wall along c_k in units of |c_k|, for links from a fluid cell to a
 *         solid cell (other entries are ignored); NULL switches the level back to halfway only.
 *   ff_fallback OAL_BFL_FALLBACK_*; q_min: links with q <= q_min or q > 1 keep halfway bounce-back
 *         (Q_MIN_THRESHOLD, 0.001).
 * The rule is applied in pull form inside the step: for link k from fluid cell x toward the wall,
 *   q < 1/2:  f_opp(k)(x, t+1) = 2q f_k*(x) + (1 - 2q) f_k*(x - c_k)
 *   q >= 1/2: f_opp(k)(x, t+1) = f_k*(x) / (2q) + (2q - 1) / (2q) f_opp(k)*(x),
 * so the reflected population arrives at x.  The reference's post-pass writes the same value into
 * x's own outgoing slot opp(k), which its pull kernel delivers to x - c_k on the next step.  With a
 * q-map, oal_force of a level sums each solid link's exchange (f_k*(x) + f_opp(k)(x, t+1)) c_k when
 * the step gathers it: the force collected in step t is that of the streaming into step t.
 * Requires the AB pattern and OAL_BB_HALFWAY.  Drops captured step graphs. */
int oal_set_bouzidi(OalPlan* plan, int level, const uint16_t* q_map, int32_t ff_fallback, float q_min);

/* Sponge layer of a level: the blend of stream_collide_cumulant! (src/collision/kernel_cumulant.jl:428-446)
 * with the per-cell strength sponge_arr (BlockLevel.sponge).  sponge [n_cells] in the reference layout
 * [x,y,z,b], values in [0, 1]; NULL (or all zero) removes it.  A cell with sp > 0 has its density and
 * velocity blended toward (1, u_inlet) before the collision uses them, rho (1 - sp) + sp and
 * u (1 - sp) + (u_inlet, 0, 0) sp; with blend_distributions 1 (SPONGE_BLEND_DISTRIBUTIONS,
 * src/io/config_loader.jl:1312) its populations are blended toward the inlet equilibrium too.  Blocks
 * holding a nonzero strength are stepped by the general kernels.  Requires the AB pattern and a level
 * without interpolated bounce-back; drops captured step graphs. */
int oal_set_sponge(OalPlan* plan, int level, const float* sponge, int32_t blend_distributions);

/* The raw level state of dump_raw_state (src/io/raw_state_dump.jl:10-47, the reference's "LBM1"
 * file, written per level as <output_dir>/raw_state/step_<NNNNNN>/level_<L>.bin), little-endian:
 *   Int32 magic 0x4C424D31; Int32 level_id, n_blocks, block_size, grid_dim_x, grid_dim_y, grid_dim_z;
 *   Float64 dx; Int32 coords[3, n_blocks] (1-based block coordinates); Float32 rho[bs,bs,bs,n_blocks];
 *   Float32 vel[bs,bs,bs,n_blocks,3]; UInt8 obstacle[bs,bs,bs,n_blocks].
 * oal_write_raw_state writes the level's current density and velocity (obstacle cells as rho = 1,
 * u = 0, as the reference's kernel leaves them); grid_dim is the largest block coordinate per axis.
 * oal_read_raw_state checks the file against the level (block count and size, coordinates, obstacle
 * mask) and starts the level at the equilibrium of the stored fields (init_eq!, src/main.jl:546-572):
 * the format keeps no non-equilibrium part.  level_id and dx receive the header's values (either may
 * be NULL).  The caller's blocks of a refined level are written and read. */
int oal_write_raw_state(OalPlan* plan, int level, int32_t level_id, double dx, const char* path);
int oal_read_raw_state(OalPlan* plan, int level, const char* path, int32_t* level_id, double* dx);

/* FP64 mass and momentum over fluid cells of each domain: out[4 * n_domains] = (mass, jx, jy, jz)
 * (src/octree_q19/src/state.jl:333-355). */
int oal_mass(OalPlan* plan, int level, double* out, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The sums of oal_mass, written
 * asynchronously on `stream` into the caller's device memory out_dev[4 * n_domains] (FP64, the same
 * values as oal_mass, bit for bit); nothing is allocated and nothing waits for the device.  The
 * per-block partials live in a fixed device array of the library (one per GPU, 2 MiB, kept until the
 * process ends), so levels of more than OAL_NOALLOC_MAX_BLOCKS blocks are rejected
 * (OAL_ERR_UNSUPPORTED).  Every call on one GPU shares that array: calls that overlap on one GPU (on
 * different streams, or from different threads) race. */
#define OAL_NOALLOC_MAX_BLOCKS 65536
int oal_mass_noalloc(OalPlan* plan, int level, double* out_dev, void* stream);

/* rho [n_cells] and u [3 * n_cells] of the current populations, FP64 host arrays in the reference
 * layout (materialize_current_state_for_io!, src/solver/solver_control.jl:830-851). */
int oal_download_macro(OalPlan* plan, int level, double* rho, double* u);

/* Stability guard (src/octree_q19/src/io_run.jl:31-65): counts fluid cells with rho <= rho_min,
 * |u| >= u_max or a non-finite population.  *bad receives the count; returns OAL_ERR_NONFINITE when
 * it is nonzero. */
int oal_guard(OalPlan* plan, int level, float rho_min, float u_max, int64_t* bad);

/* Waits for all work of the plan. */
int oal_sync(OalPlan* plan);

/* Times n steps of one level on the plan's stream with HIP events: *ms_total is the time of the n
 * steps (all their launches), *ms_kernel the mean duration of the bulk launches of one step (the
 * regular and interior block lists, events on the plan's stream around them; the mean step time
 * when the level has no such blocks).  level OAL_ALL_LEVELS times n steps of oal_step (time steps
 * 0 .. n-1, every level, transfers and graph replay included); *ms_kernel is then the mean step time. */
#define OAL_ALL_LEVELS (-1)
int oal_time_steps(OalPlan* plan, int level, int n, int flags, double* ms_total, double* ms_kernel);

/* Replaces: nothing.  Bandwidth of a streaming 16-byte device copy of `bytes` (read + write counted),
 * mean over `reps` back-to-back launches after one warm-up: the measured HBM ceiling beside the
 * 8 TB/s datasheet peak. */
int oal_copy_bandwidth(int device, int64_t bytes, int reps, double* gbs);

/* Launch shape of the step kernels.  pipe_planes: z-planes of a 16^3 block one workgroup of the
 * interior kernel walks with its next plane's loads in flight (1 = one plane per workgroup, no
 * pipelining; 2, 4, 8 or 16).  overlap: 1 = boundary and face launches on a second stream,
 * concurrent with the bulk launch; 0 = every launch in order on the plan's stream.  face_kernel
 * (pull/AB pattern, blocks on a domain face): 0 = the pipelined kernel applies the domain-face rules
 * itself; 1 = it skips the cells on the faces and a separate face-cell kernel steps them.  graph:
 * 1 = oal_step on the plan's own stream (stream NULL) replays a hipGraph of one (coarse) step,
 * captured on first use for each arrangement of the levels' buffers, with the step index read from
 * device memory (a step that collects the force runs as ordinary launches); 0 = the host issues
 * every launch.  Refined plans start with graph 1, one-level plans with 0.  Setting the tuning drops
 * the captured graphs. */
typedef struct {
    int32_t pipe_planes;
    int32_t overlap;
    int32_t face_kernel;
    int32_t graph;
    int32_t pair_kernel;   /* FP32 and 16-bit storage, AB pattern, pipe_planes > 1, face_kernel 0, no WALE: the
                              regular blocks step two cells per lane (8-byte FP32 / 4-byte 16-bit words holding
                              both): 1 in one launch (the FP32 default), 2 with the blocks on a domain face in a
                              launch of their own (the 16-bit default); 0 = one cell per lane; the same bits
                              every way */
    int32_t reserved[3];
} OalTuning;

/* number of launch shapes the autotuner times: pipe_planes {1, 2, 4, 8, 16} x overlap {1, 0} x
 * face_kernel {0, 1} */
#define OAL_TUNE_SHAPES 20

/* Replaces: nothing (the reference picks its launch shape statically, src/collision/timestep.jl:1479). */
int oal_get_tuning(const OalPlan* plan, OalTuning* tuning);
int oal_set_tuning(OalPlan* plan, const OalTuning* tuning);

/* The project's autotuner: times every candidate launch shape on `level` for n_steps steps each,
 * keeps the fastest and restores the populations the plan held before; *best receives the choice.
 * ms_per_step[OAL_TUNE_SHAPES] (may be NULL) receives the mean step time of shape i: pipe_planes
 * {1, 2, 4, 8, 16}[i % 5], overlap 1 for (i / 5) even, face_kernel 1 for i >= 10. */
int oal_autotune(OalPlan* plan, int level, int n_steps, OalTuning* best, double* ms_per_step);

#ifdef __cplusplus
}
#endif

#endif /* OAL_H */
