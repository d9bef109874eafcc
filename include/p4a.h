/*
 * p4a.h - C ABI of libp4a.so, an MI355X (gfx950) implementation of the
 * Particles4All simulation step: Position Based Fluids with rigid bodies in
 * one constraint loop, XSPH viscosity, Akinci surface tension and the
 * anisotropic-kernel covariance pass.
 *
 * Every entry point replaces one member of the reference's `class Sim`
 * (src/sim.js:9-814) or one host routine it calls; each declaration cites the
 * reference code it stands in for.  Conventions:
 *   - every function returns int status, 0 (P4A_OK) on success;
 *   - p4a_last_error(h) returns a message owned by the library (valid until
 *     the next call on that handle);
 *   - a handle is used from one host thread at a time; GPU work is queued on
 *     the handle's own HIP stream and p4a_sync() waits for it;
 *   - all pointers are plain host pointers unless the name says `device`;
 *   - inputs the reference clamps or ignores are clamped or ignored here too
 *     (appendFluid past capacity, holdBody with an unknown body, a ray
 *     impulse with radius <= 0, a resize to the same box).
 */
#ifndef P4A_H
#define P4A_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define P4A_API __attribute__((visibility("default")))

typedef struct p4a_sim p4a_sim;          /* one scene on one GPU */
typedef struct p4a_scene_data p4a_scene_data; /* host-side scene arrays */

enum p4a_status {
  P4A_OK = 0,
  P4A_ERR_ARG = 1,     /* null pointer, negative size, unknown enum */
  P4A_ERR_HIP = 2,     /* a HIP runtime call failed (message says which) */
  P4A_ERR_STATE = 3,   /* call needs reset() first, or similar */
  P4A_ERR_NOMEM = 4,   /* device or host allocation failed */
  P4A_ERR_RANGE = 5,   /* offset/size outside the buffer */
  P4A_ERR_NODEV = 6    /* no HIP device with that ordinal */
};

/* Arithmetic form of the neighbour kernels. */
enum p4a_precision {
  P4A_PREC_FP32 = 0,         /* float positions, float sums (reference form) */
  P4A_PREC_FP16_OFFSET = 1   /* solver iterations (lambda, delta) read
                                neighbours as 8-byte records: x as fp32,
                                y and z as fp16 offsets from the centre of
                                the cell each particle was binned into; fp32
                                lambda, sums and master positions.  Fast
                                math only: with P4A_MATH_EXACT, reset and
                                set_params fail with P4A_ERR_ARG */
};

/* Rounding of sqrt and division inside the neighbour kernels. */
enum p4a_math {
  P4A_MATH_FAST = 0,   /* hardware rsq/rcp (1 ulp): production form */
  P4A_MATH_EXACT = 1   /* IEEE-rounded sqrt and division, no contraction: the
                          results equal an op-by-op float evaluation */
};

/*
 * Simulation parameters: defaultParams (src/scene.js:11-39) plus the URL
 * flags that reach Sim (src/main.js:361-390).  FP64 host values; the
 * library derives the float uniforms exactly as Sim.uploadParams does
 * (src/sim.js:371-413).
 */
typedef struct p4a_params {
  double spacing;            /* particle spacing d; h = 2d (scene.js:160-161) */
  double rest_density;       /* rho0 */
  double gravity;            /* g, acting along -y */
  int32_t substeps;          /* per 1/60 s frame (sim.js:418-419) */
  int32_t iterations;        /* PBF solver iterations I */
  double cfm_epsilon_rel;    /* CFM relaxation relative to denomRest */
  double scorr_k;            /* tensile-instability correction k */
  double scorr_dq;           /* dq as a fraction of h */
  double xsph_c;             /* XSPH viscosity c */
  double omega;              /* over-relaxation of delta-p */
  int32_t sor_average;       /* 1: divide delta-p by neighbour count */
  double surface_tension_k;  /* Akinci gamma; 0 disables normals+tension */
  double box[3];             /* box size; boxMin is the origin */
  int32_t no_boundary;       /* 1: ignore the wall samples (boundary=0) */
  int32_t pour_budget;       /* extra capacity for appendFluid; -1 = nFluid
                                (the reference's default, sim.js:94) */
  int32_t precision;         /* enum p4a_precision */
  int32_t math;              /* enum p4a_math */
  int32_t use_graph;         /* 1: replay each substep as a captured hipGraph */
  double neighbour_skin;     /* FP32 forms: skin S of the per-substep neighbour
                                masks, in units of h (0 <= S <= 1; 0: every
                                solver kernel walks the full 27-cell stencil).
                                The results are the full walk's bit for bit
                                either way.  No counterpart in the reference. */
  int32_t neighbour_min_particles; /* the masks are used for scenes of at
                                least this many particles (default 524288:
                                below it their extra launches cost more than
                                the candidates they save); >= 0.  No
                                counterpart in the reference. */
} p4a_params;

/* Fill `p` with defaultParams('small') (src/scene.js:14-31). */
P4A_API int p4a_params_default(p4a_params* p);

/* ---------------------------------------------------------------------
 * Scene construction (host side).  buildScene (src/scene.js:159-315)
 * and its helpers fluidBlock (43-64), sampleBody (137-157) and
 * boundaryParticles (76-135), plus lattice scenes for synthetic workloads.
 * ------------------------------------------------------------------- */

enum p4a_body_shape {
  P4A_SHAPE_SPHERE = 0,     /* sampleBody 'sphere' (scene.js:148) */
  P4A_SHAPE_TORUS = 1,      /* sampleBody 'torus' (scene.js:149-151) */
  P4A_SHAPE_BOX = 2,        /* sampleBody box: half extents .9s/.7s/.9s */
  P4A_SHAPE_LATTICE_CUBE = 3,   /* `size` particles per edge */
  P4A_SHAPE_LATTICE_SPHERE = 4  /* lattice sites with |o| <= size spacings */
};

typedef struct p4a_body_spec {
  int32_t shape;         /* enum p4a_body_shape */
  double size;           /* lattice scenes: edge (cube) or radius (sphere) in
                            spacings; presets use body_size for every body */
  double density;        /* body density relative to the fluid; presets map
                            0 or NaN to 0.5 (scene.js:169) */
  double start_y;        /* presets: centre height / box y; 0 or NaN = 0.75 */
  double centre[3];      /* lattice scenes only; presets place bodies */
} p4a_body_spec;

enum p4a_scene_kind {
  P4A_SCENE_PRESET = 0,  /* buildScene(p) exactly as the reference */
  P4A_SCENE_LATTICE = 1  /* jittered lattice block + explicit bodies */
};

typedef struct p4a_scene_spec {
  int32_t kind;                  /* enum p4a_scene_kind */
  /* preset scenes (scene.js:11-39, 159-315) */
  int32_t target_particle_count;
  double body_size;              /* 0: 0.12 * min(box) (scene.js:172) */
  /* lattice scenes */
  int32_t lattice[3];            /* sites per axis */
  double origin[3];              /* block min corner; site = origin+(i+.5)d */
  double jitter;                 /* uniform +-jitter*d per coordinate */
  uint32_t seed;                 /* LCG seed (scene.js:52-56 idiom) */
  int32_t calibrate_unjittered;  /* 1: m and denomRest from the unjittered
                                    lattice; 0: as buildScene (jittered max) */
  /* bodies (both kinds) */
  int32_t n_bodies;
  const p4a_body_spec* bodies;
} p4a_scene_spec;

/* A view of built scene arrays, laid out as buildScene's return value
 * (src/scene.js:313-314).  Pointers stay owned by the p4a_scene_data. */
typedef struct p4a_scene {
  int32_t n, n_fluid, n_body;
  const float* pos;        /* n*4, w = 0 */
  const float* vel;        /* n*4 */
  const float* rest;       /* n*4, body-local offset from the body centroid */
  const uint32_t* phase;   /* n*4, x = 1-based body id, 0 = fluid */
  int32_t n_boundary;
  const float* boundary_pts;  /* nb*3 */
  const float* boundary_psi;  /* nb */
  int32_t n_bodies;
  const double* body_centre;  /* nb*3 */
  const double* body_density; /* per body, relative to the fluid */
  const int32_t* body_count;  /* particles per body */
  const int32_t* body_spec;   /* index into the spec's body list (bodies
                                 that sample no particle are dropped,
                                 scene.js:183) */
  double mass, h, d, denom_rest;
} p4a_scene;

/* buildScene (scene.js:159-315) for P4A_SCENE_PRESET (params->box,
 * spacing, rest_density and spec->target_particle_count, body list),
 * or a jittered lattice scene for P4A_SCENE_LATTICE. */
P4A_API int p4a_scene_build(const p4a_params* p, const p4a_scene_spec* spec,
                            p4a_scene_data** out);
P4A_API int p4a_scene_view(const p4a_scene_data* s, p4a_scene* view);
P4A_API void p4a_scene_free(p4a_scene_data* s);

/* boundaryParticles (scene.js:76-135): wall samples of the box shell and
 * their Akinci volumes psi.  pts (nb*3) and psi (nb) are written when
 * non-null; *count always.  Call once with nulls to size the arrays. */
P4A_API int p4a_boundary_build(const double box[3], double d, double h,
                               double rest_density, float* pts, float* psi,
                               int32_t capacity, int32_t* count);

/* ---------------------------------------------------------------------
 * The Sim surface (src/sim.js)
 * ------------------------------------------------------------------- */

/* new Sim(device) (sim.js:10-18): bind to HIP device `device`. */
P4A_API int p4a_create(int device, p4a_sim** out);
/* Frees every device buffer of the handle. */
P4A_API int p4a_destroy(p4a_sim* h);
P4A_API const char* p4a_last_error(const p4a_sim* h);

/* Sim.reset(params) (sim.js:66-221) with the scene already built (the
 * reference calls buildScene inside reset; callers here pass its result).
 * Allocates for cap = n + pour budget, sorts the wall samples
 * (sortBoundary, sim.js:255-286), uploads the state and primes the grid
 * (primeGrid, sim.js:223-253). */
P4A_API int p4a_reset(p4a_sim* h, const p4a_params* p, const p4a_scene* scene);

/* Re-read the live parameters that uploadParams takes each step
 * (substeps, iterations, omega, xsph, scorr, cfm, tension, gravity;
 * sim.js:371-413).  Geometry fields (box, spacing) are ignored here. */
P4A_API int p4a_set_params(p4a_sim* h, const p4a_params* p);

/* Sim.step(frameDt) (sim.js:415-656): time bank, then `sub` substeps of
 * predict, grid build, bodies, I solver iterations and the finish kernels,
 * then the stats reduction.  *substeps_run may be null. */
P4A_API int p4a_step(p4a_sim* h, double frame_dt, int32_t* substeps_run);

/* Run exactly `count` substeps of length dt (no time bank; for benches). */
P4A_API int p4a_substeps(p4a_sim* h, double dt, int32_t count);

/* Sim.resizeBox(newBox) (sim.js:687-772). */
P4A_API int p4a_resize_box(p4a_sim* h, const double box[3]);

/* Sim.appendFluid(pos, vel) (sim.js:774-809): pos3/vel3 are count*3
 * floats; *appended receives the number actually added (0 when full).
 * Non-finite positions or velocities are refused (P4A_ERR_ARG, nothing
 * appended). */
P4A_API int p4a_append_fluid(p4a_sim* h, const float* pos3, const float* vel3,
                             int32_t count, int32_t* appended);

/* Sim.applyRayImpulse(origin, dir, impulse, radius, speedLimit)
 * (sim.js:658-676).  Non-finite arguments are refused (P4A_ERR_ARG). */
P4A_API int p4a_apply_ray_impulse(p4a_sim* h, const double origin[3],
                                  const double dir[3], const double impulse[3],
                                  double radius, double speed_limit);

/* Sim.holdBody / Sim.releaseBody (sim.js:678-685).  A non-finite target,
 * rate or limit is refused (P4A_ERR_ARG, the hold unchanged). */
P4A_API int p4a_hold_body(p4a_sim* h, int32_t body_1based, const double target[3],
                          double rate, double limit);
P4A_API int p4a_release_body(p4a_sim* h);

/* Buffers readable through p4a_read / p4a_device_ptr.  Particle arrays are
 * in the live (cell-sorted) order of the last substep; P4A_BUF_PID maps a
 * live slot to the particle's id (its index in the scene arrays or the
 * append order), which the reference has no counterpart of. */
enum p4a_buf {
  P4A_BUF_POS = 0,         /* float4 n   (sim.buf.posA/B, livePos) */
  P4A_BUF_VEL = 1,         /* float4 n   (liveVel) */
  P4A_BUF_PRED = 2,        /* float4 n   (pred of the live parity) */
  P4A_BUF_BODY = 3,        /* uint4 n, x = body id (liveBody); p4a_read builds
                              it on the host; p4a_device_ptr makes it a device
                              buffer kept current after every step, stage and
                              append from then on (the SSFR renderer binds it
                              with P4A_BUF_SMOOTH_POS and P4A_BUF_ANISO_G) */
  P4A_BUF_REST = 4,        /* float4 n   (restA/B) */
  P4A_BUF_LAMBDA = 5,      /* float n */
  P4A_BUF_DENSITY = 6,     /* float n */
  P4A_BUF_SLOT = 7,        /* uint n: slot of the previous order's i (scatterSlot) */
  P4A_BUF_CORR = 8,        /* float4 n */
  P4A_BUF_NORMAL = 9,      /* float4 n */
  P4A_BUF_CELL_START = 10, /* uint nCells+2 */
  P4A_BUF_BPOS = 11,       /* float4 nb, sorted wall samples, w = psi */
  P4A_BUF_BCELL_START = 12,/* uint nCells+2 */
  P4A_BUF_BODY_CENTRE = 13,/* float4 nBodies (bodyCentre) */
  P4A_BUF_BODY_ROT = 14,   /* float4 3*nBodies, rows of R (bodyRot) */
  P4A_BUF_PID = 15,        /* uint n */
  P4A_BUF_KEY = 16,        /* uint n: cell id of each particle, pre-sort order */
  P4A_BUF_ORDER = 17,      /* uint n: pre-sort index that landed in each slot */
  P4A_BUF_BPSI = 18,       /* float nb */
  P4A_BUF_SMOOTH_POS = 19, /* float4 n: smoothed centres of the last p4a_aniso */
  P4A_BUF_ANISO_G = 20,    /* float4 2n: packed G of the last p4a_aniso */
  P4A_BUF_OFS = 21,        /* 8 bytes n: FP16-offset records of the current
                              pred (float x; half y, z offsets from the
                              binned cell's centre) */
  P4A_BUF_CELL_OF_SLOT = 22, /* uint n: cell each slot was binned into */
  P4A_BUF_DRIFT = 23,      /* float 256: per position update since the last grid
                              build (delta, shape matching, in launch order),
                              the largest per-axis displacement of a particle */
  P4A_BUF_FIELD = 24,      /* float n_verts: the last p4a_field_generate,
                              after any p4a_field_smooth */
  P4A_BUF_FIELD_NORMAL = 25, /* float n_verts: the last p4a_mesh_normal_field */
  P4A_BUF_MESH_VERTS = 26, /* uint2 3*max_triangles: packed triangle vertices
                              of the last p4a_mesh_march */
  P4A_BUF_MESH_INDIRECT = 27, /* uint 4: draw-indirect arguments of the last
                              march (vertex count, 1, 0, 0) */
  P4A_BUF_BODY_PACKED = 28, /* float4 6*n_bodies: the last p4a_body_pack */
  P4A_BUF_NEIGHBOUR_STATE = 29, /* uint 1025: largest travel since the last
                              neighbour-mask build (float bits), then the
                              targets the last two list launches left to the
                              full walk (256 counts per launch parity), then
                              the targets queued as having left their build
                              cells (256 counts per solver iteration parity).
                              Read only; no counterpart in the reference */
  P4A_BUF_COUNT_
};

/* Copy `bytes` bytes at byte `offset` of buffer `id` to host `dst`
 * (synchronous; the reference's __readBuf, main.js:810-820). */
P4A_API int p4a_read(p4a_sim* h, int32_t id, void* dst, size_t offset, size_t bytes);
/* Overwrite part of a state buffer from the host (test and restore use).
 * The index buffers (PID, ORDER, KEY, SLOT, CELL_START, BCELL_START,
 * CELL_OF_SLOT) and the synthesised BODY/REST views are not writable;
 * POS, VEL and PRED take whole finite floats only (P4A_ERR_ARG). */
P4A_API int p4a_write(p4a_sim* h, int32_t id, const void* src, size_t offset, size_t bytes);
/* Device address and size of a buffer for zero-copy interop. */
P4A_API int p4a_device_ptr(p4a_sim* h, int32_t id, void** ptr, size_t* bytes);

/* Scalars the reference exposes as Sim fields (sim.js:71-221, 415-429). */
typedef struct p4a_info {
  int32_t n, cap, n_fluid, n_body_parts, n_bodies, n_boundary, n_cells;
  int32_t grid_dim[3];
  int32_t parity, gen, last_substeps, pour_budget;
  double h, mass, denom_rest, sim_time, last_advanced, time_bank;
  double body_scale, cov_scale;
} p4a_info;
P4A_API int p4a_get_info(p4a_sim* h, p4a_info* out);

/* Sim.stats after the last step (sim.js:619-633): avgRho, maxRho (both
 * relative to rho0), maxSpeed, kinetic energy; reduced exactly in FP64. */
P4A_API int p4a_stats(p4a_sim* h, double out[4]);

/* The stats reduction of p4a_stats on the current live state (densities of
 * the last lambda, live velocities), written to caller device memory
 * `out_dev` (4 doubles: avgRho, maxRho, maxSpeed, kinetic energy) and
 * enqueued on the handle's stream without waiting.  It allocates nothing:
 * its per-block partial sums live in a fixed device array of the library
 * (512 KiB on each GPU it runs on, held until the process ends), so callers
 * that cannot allocate (code under HIP graph capture, serving loops) can
 * call it.  No counterpart in the reference.  Handles with more than
 * P4A_STATS_NOALLOC_MAX_PARTICLES particles are refused (P4A_ERR_RANGE).
 * Calls on one GPU share the array: calls that overlap on one GPU (two
 * streams at once) race. */
#define P4A_STATS_NOALLOC_MAX_PARTICLES (1u << 25)
P4A_API int p4a_stats_noalloc(p4a_sim* h, double* out_dev);

/* Sim.bodyPose (sim.js:635-655): centre nBodies*3, rot nBodies*9
 * (column-major, as the reference's p.rot). */
P4A_API int p4a_body_pose(p4a_sim* h, float* centre, float* rot);

/* ---------------------------------------------------------------------
 * Anisotropic-kernel pass (src/aniso_wgsl.js:3-159, parameters as
 * SurfaceMesh.buildAnisotropy fills them, src/mesh.js:207-261): per live
 * particle, the weighted neighbour mean and covariance within
 * R = radius_scale*h over the current grid, Laplacian-smoothed centre, a 3x3
 * Jacobi eigensolve and the clamped ellipsoid matrix G.  Output buffers:
 * P4A_BUF_SMOOTH_POS (xyz, 0) and P4A_BUF_ANISO_G, two float4 per particle:
 * (G00, G01, G02, G11), (G12, G22, det G, neighbour count incl. self).
 * ------------------------------------------------------------------- */
typedef struct p4a_aniso_params {
  double radius_scale;   /* anisoRadiusScale: R = radius_scale * h, <= 2 */
  double lambda;         /* anisoLambda: centre smoothing */
  double kr;             /* anisoRatio: largest/smallest axis ratio cap */
  double kn;             /* anisoLonely: isotropic scale of lonely particles */
  double ks;             /* anisoKs, divided by 0.15 R^2 as mesh.js:226 */
  double stretch;        /* anisoStretch: cap on the largest extent */
  double min_extent;     /* lower bound of each extent (0 = none) */
  int32_t neps;          /* anisoMinNeighbours: blend window centre */
  int32_t fp64_fallback; /* 1: re-solve in FP64 when the eigenvalue gap is
                            below gap_tol * max|eigenvalue| */
  double gap_tol;        /* 1e-4 */
} p4a_aniso_params;

/* SurfaceMesh defaults (mesh.js:81-87) with the FP64 fallback on. */
P4A_API int p4a_aniso_params_default(p4a_aniso_params* a);
/* Run the pass on the live positions; *fp64_solves (may be null) receives
 * the number of particles whose eigensolve took the FP64 path. */
P4A_API int p4a_aniso(p4a_sim* h, const p4a_aniso_params* a, int32_t* fp64_solves);

/* ---------------------------------------------------------------------
 * Density field on a voxel-vertex lattice.
 *
 * p4a_field_configure = SurfaceMesh.configure (src/mesh.js:96-128):
 * voxel = longest(box_max - box_min) / resolution, vert_dim = ceil(size /
 * voxel) + 3 per axis, origin = box_min - voxel; vertex v sits at origin +
 * voxel * v; flat index (v.z * vert_dim.y + v.y) * vert_dim.x + v.x.
 * max_triangles is the march's triangle budget (<= 0: the reference's
 * default 3,000,000; at most 2^30).
 *
 * p4a_field_generate = SurfaceMesh.generateField (src/mesh.js:263-271):
 *  P4A_FIELD_ISOTROPIC: meshDensityWGSL (src/mesh_wgsl.js:23-69), per
 *    vertex p the sum over the 27 cells of the (stale) simulation grid
 *    around p of m / max(rho_j, rho0/2) * poly6 * (h^2 - r^2)^3, r^2 < h^2,
 *    over the live positions;
 *  P4A_FIELD_ANISOTROPIC: the Yu-Turk anisotropic field of the last
 *    p4a_aniso pass (no reference counterpart): sum of m / max(rho_j,
 *    rho0/2) * 315/(64 pi) * det(G_j) * (1 - |G_j (p - smoothPos_j)|^2)^3
 *    inside the ellipsoid |G_j (p - smoothPos_j)| < 1.  Terms are rounded to
 *    multiples of 2^-40 and summed in 64-bit integers, so the result does
 *    not depend on the order of the GPU's atomics.
 * skip_bodies: 1 skips body particles, 0 keeps them, -1 skips them when
 * the scene has bodies (mesh.js:145).  Read the result with P4A_BUF_FIELD.
 * ------------------------------------------------------------------- */
enum p4a_field_kind {
  P4A_FIELD_ISOTROPIC = 0,
  P4A_FIELD_ANISOTROPIC = 1
};
typedef struct p4a_field_info {
  int32_t vert_dim[3];
  int32_t pad_;
  double voxel;
  double origin[3];
  int64_t n_verts;
  int64_t max_triangles;
} p4a_field_info;
P4A_API int p4a_field_configure(p4a_sim* h, const double box_min[3], const double box_max[3],
                                int32_t resolution, int64_t max_triangles);
P4A_API int p4a_field_generate(p4a_sim* h, int32_t kind, int32_t skip_bodies);
P4A_API int p4a_field_get_info(p4a_sim* h, p4a_field_info* out);

/* ---------------------------------------------------------------------
 * Surface extraction over the field lattice.
 *
 * p4a_field_smooth = SurfaceMesh.smoothField (src/mesh.js:273-291): `passes`
 * rounds of the separable blur of meshSmoothWGSL (mesh_wgsl.js:70-91),
 * dst(v) = 0.25 S(v - a) + 0.5 S(v) + 0.25 S(v + a) along a = x, y, z, S
 * clamping v into the lattice; in place on P4A_BUF_FIELD.
 *
 * p4a_mesh_normal_field = SurfaceMesh.buildNormalField (mesh.js:293-313):
 * the same blur of the field into P4A_BUF_FIELD_NORMAL, which the march
 * then differentiates for normals; passes <= 0 makes the march use the
 * field itself (normalFieldReady = false).
 *
 * p4a_mesh_march = the march of SurfaceMesh.generate (mesh.js:320-331,
 * meshMarchWGSL mesh_wgsl.js:93-233, meshIndirectWGSL 235-244): marching
 * tetrahedra (six per voxel around its 0-7 diagonal), corners inside when
 * field > iso, vertices on the crossed edges by linear interpolation,
 * normals -grad of the normal field by central differences.  Each vertex
 * is packed into two words as packVertex does: x | y << 16, z | (octahedral
 * normal, 8 + 8 bits) << 16, positions in lattice units / (vert_dim - 1)
 * quantised to 16 bits.  Triangles are numbered in voxel order (flat voxel
 * index, then tetrahedron, then triangle), so the vertex buffer is the
 * same on every run (the reference appends with an atomic counter, in no
 * fixed order); triangles numbered >= max_triangles are dropped.
 * Asynchronous; p4a_mesh_triangles waits and returns the total count
 * (dropped triangles included, as the reference's counter) and the number
 * written.
 * ------------------------------------------------------------------- */
P4A_API int p4a_field_smooth(p4a_sim* h, int32_t passes);
P4A_API int p4a_mesh_normal_field(p4a_sim* h, int32_t passes);
P4A_API int p4a_mesh_march(p4a_sim* h, float iso);
P4A_API int p4a_mesh_triangles(p4a_sim* h, int64_t* total, int64_t* emitted);

/* ---------------------------------------------------------------------
 * Body poses for the solid renderer: bodyPackWGSL (src/solid_wgsl.js:3-26)
 * as Solids.build / Solids.pack run it (src/solids.js:107-161).  stat holds
 * 8 floats per body as Solids.build fills it (half extents xyz, shape id;
 * colour rgb, 0); null keeps the last upload.  count must equal n_bodies.
 * Writes P4A_BUF_BODY_PACKED from the current poses, per body 6 float4:
 * (centre, shape id), (half extents, 0), the three columns of R (each
 * (R0k, R1k, R2k, 0)), (colour, 0).  Asynchronous.
 * ------------------------------------------------------------------- */
P4A_API int p4a_body_pack(p4a_sim* h, const float* stat, int32_t count);

/* ---------------------------------------------------------------------
 * State snapshot (no reference counterpart: the reference keeps its state
 * only on the GPU).  For reproducible long runs and regression fixtures:
 * a sim loaded from a snapshot continues bit for bit as the saved one
 * would have (same steps, same results), because the snapshot holds the
 * live particle order, the FP64 body state, the last grid, the time bank
 * and the parameters.  p4a_snapshot_size gives the bytes save writes;
 * p4a_snapshot_load resets the handle from a snapshot (any handle: scene,
 * capacity and box come from the snapshot).  Native byte order; a header
 * records the layout version and sizes, and load rejects a snapshot of
 * another version, a truncated one or one whose ids are not a
 * permutation.  Render-side state (anisotropy, fields, meshes) is not
 * part of it.
 * ------------------------------------------------------------------- */
P4A_API int p4a_snapshot_size(p4a_sim* h, size_t* bytes);
P4A_API int p4a_snapshot_save(p4a_sim* h, void* dst, size_t bytes);
P4A_API int p4a_snapshot_load(p4a_sim* h, const void* src, size_t bytes);

/* Wait for all queued GPU work of the handle. */
P4A_API int p4a_sync(p4a_sim* h);

/* ---------------------------------------------------------------------
 * Stage-level entry points: run one kernel group of a substep on the
 * current state, so each can be checked against the oracle on identical
 * inputs.  The stages follow the reference's pass order (sim.js:458-596).
 * ------------------------------------------------------------------- */
enum p4a_stage {
  P4A_STAGE_PREDICT = 0,    /* bodyDrag (if held) + predict (wgsl.js:56-70) */
  P4A_STAGE_GRID = 1,       /* count, scan, stable scatter (wgsl.js:352-475) */
  P4A_STAGE_BODY_PRE = 2,   /* body centroids (wgsl.js:628-668, 174-185) */
  P4A_STAGE_LAMBDA = 3,     /* wgsl.js:477-543 */
  P4A_STAGE_DELTA = 4,      /* wgsl.js:545-612 */
  P4A_STAGE_BODY_SOLVE = 5, /* shape matching (wgsl.js:628-783) */
  P4A_STAGE_FINISH = 6,     /* velFromPos, xsph, normals, tension, finalize */
  P4A_STAGE_STATS = 7       /* wgsl.js:785-832 */
};
P4A_API int p4a_run_stage(p4a_sim* h, int32_t stage, double dt);

/* Per particle (live order): number of fluid neighbours with r^2 < h^2 in
 * the 27-cell stencil of the current grid, and two checksums of their
 * indices (sum of j, sum of j*j mod 2^32).  Integer outputs, n each. */
P4A_API int p4a_neighbour_census(p4a_sim* h, uint32_t* count, uint32_t* sum_j,
                                 uint32_t* sum_j2);

/* Per-kernel timing of the last p4a_substeps call: names and mean ms per
 * launch, measured with HIP events on the handle's stream.  `names` gets
 * a newline-separated list.  Enabled by p4a_set_timing(h, 1). */
P4A_API int p4a_set_timing(p4a_sim* h, int32_t on);
P4A_API int p4a_timing(p4a_sim* h, char* names, size_t names_bytes, double* ms,
                       int32_t* launches, int32_t capacity, int32_t* count);

/* Device copy probe: copies `bytes` (a multiple of 16) between two fresh
 * device buffers `reps` times with the library's float4 copy kernel and
 * returns the mean ms per copy.  Calibrates hardware byte counters and
 * gives the attainable streaming bandwidth; no counterpart in the
 * reference. */
P4A_API int p4a_copy_probe(int device, size_t bytes, int32_t reps, double* ms);

/* Autotuner: times the neighbour kernels' tile shapes on the live scene
 * (lambda for the one-record kernels; xsph+normals followed by tension for
 * the two-record ones, as a substep runs them) and keeps the fastest for
 * later steps.  In the FP16-offset form group 1 times that form's lambda,
 * which takes only the thread count of the direct-walk shapes 0-2; the
 * other group-1 shapes report -1 (P4A_ERR_ARG when the FP16 records are
 * stale).  chosen[0..1] receive the variant ids; ms (capacity entries) the
 * mean ms of every variant tried, group 1 first.  The particle state is
 * not changed. */
P4A_API int p4a_tune(p4a_sim* h, int32_t chosen[2], double* ms, int32_t capacity);
/* Name of tile variant v of group 1 or 2. */
P4A_API const char* p4a_tile_variant_name(int32_t group, int32_t v);

/* Library version string. */
P4A_API const char* p4a_version(void);

/* ---------------------------------------------------------------------
 * Batched scenes (no counterpart in the reference, which runs one scene
 * per Sim): many independent scenes stepped together.  Every stage of a
 * substep is one launch over all scenes (particles, cells and wall samples
 * of all scenes concatenated; each workgroup lies in one scene and reads
 * that scene's parameter block), and the substep is one hipGraph.  Scenes
 * keep their own box, h, mass, gravity, XSPH, SOR and solver constants;
 * they must share the iteration count, tension on/off and the math mode,
 * and run the FP32 form.  Scene arrays are copied; no pour capacity.
 * When p4a_batch_create fails after its argument checks, *out still holds a
 * handle whose p4a_batch_last_error names the cause; destroy it as usual.
 * ------------------------------------------------------------------- */
typedef struct p4a_batch p4a_batch;
P4A_API int p4a_batch_create(int32_t device, int32_t n_scenes, const p4a_params* params,
                             const p4a_scene* scenes, p4a_batch** out);
P4A_API int p4a_batch_destroy(p4a_batch* b);
P4A_API const char* p4a_batch_last_error(p4a_batch* b);
/* count substeps of dt for every scene, then the per-scene stats */
P4A_API int p4a_batch_substeps(p4a_batch* b, double dt, int32_t count);
/* Autotuner of the batch, as p4a_tune: times the traversal shapes that have
 * a batched form (variant ids 0, 1, 2, 5, 6, 7, 8, 9) on the batch's current
 * state with substep dt and keeps the fastest; chosen[0..1] receive them,
 * ms (capacity entries) the mean ms of each, group 1 first.  In exact math
 * every shape gives the same bits; in fast math the one-phase walks and the
 * two-phase and staged walks (which test r^2 < h^2 before the term) differ
 * in the last bits. */
P4A_API int p4a_batch_tune(p4a_batch* b, double dt, int32_t chosen[2], double* ms,
                           int32_t capacity);
P4A_API int p4a_batch_sync(p4a_batch* b);
P4A_API int p4a_batch_scene_size(p4a_batch* b, int32_t scene, int32_t* n);
/* One scene's slice (live order) of P4A_BUF_POS, VEL, PRED, DENSITY, LAMBDA
 * or PID (ids local to the scene). */
P4A_API int p4a_batch_read(p4a_batch* b, int32_t scene, int32_t buf, void* dst, size_t offset,
                           size_t bytes);
/* p4a_stats of one scene after the last p4a_batch_substeps. */
P4A_API int p4a_batch_stats(p4a_batch* b, int32_t scene, double out[4]);

#ifdef __cplusplus
}
#endif
#endif /* P4A_H */
