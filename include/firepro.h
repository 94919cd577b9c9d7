/*
 * firepro.h — C ABI of the MI355X-native sparse-brick fire/smoke solver.
 *
 * The entry points replace the compute passes of Fire Pro's engine layer
 * (src/engine/FluidSimulation.ts) behind the members that the library layer
 * (src/library/FireSimulation.ts) calls. Every structure is plain data: numbers,
 * fixed arrays and borrowed pointers, read during the call only. Host-side geometry
 * (positions, sizes) is double, as JavaScript numbers are; the solver converts it to
 * 32-bit floats where the reference packs Float32Arrays for the GPU.
 *
 * Status: every entry point returns FP_OK (0) or a negative FP_E* code; the message
 * of the last failure on the calling thread is fp_last_error().
 */
#ifndef FIREPRO_H
#define FIREPRO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FP_OK 0
#define FP_EINVAL (-1)   /* invalid argument: the reference throws Error here */
#define FP_EDEVICE (-2)  /* HIP runtime failure */
#define FP_ESTATE (-3)   /* call out of order (released handle, missing step) */
#define FP_ENOMEM (-4)   /* device or host allocation failed */

#define FP_SHAPE_SPHERE 0
#define FP_SHAPE_BOX 1
#define FP_SHAPE_MESH 2 /* sources only: emission from a mesh surface (MeshSourceRasterizer.ts) */

#define FP_FORCE_WIND 0
#define FP_FORCE_TURBULENCE 1
#define FP_FORCE_VORTEX 2
#define FP_FORCE_RADIAL 3

#define FP_MAX_FLAME_KNOTS 8

/* Pressure policies (fp_sim_set_pressure). */
#define FP_PRESSURE_VCYCLE 0 /* the reference: one cold V-cycle, FP32 (PressureSolver.ts:371-402) */
#define FP_PRESSURE_FP32 1   /* defect correction, V-cycle levels stored in FP32 */
#define FP_PRESSURE_BF16 2   /* defect correction, V-cycle levels stored in BF16 */
#define FP_PRESSURE_FP16 3   /* defect correction, V-cycle levels stored in FP16 */

/* SimulationGrid (src/engine/types.ts:26-36). Construction-only. */
typedef struct fp_grid {
  double voxel_size;        /* meters, 0.0001..16 (FluidSimulation.ts:362-363) */
  int32_t velocity_divisor; /* 1, 2 or 4 */
  int32_t brick_size;       /* 8, 16 or 32 */
  int32_t smoke_divisor;    /* 1, 2 or 4: smoke on its own grid, divisor x voxel wide (smoke.wgsl);
                               2 and 4 are not batched (fp_batch_step) */
  int32_t scalar_maccormack;/* 1: bounded MacCormack scalars; 0: first order */
} fp_grid;

/* SimulationSettings (src/engine/types.ts:235-243). */
typedef struct fp_settings {
  double buoyancy, smoke_weight, velocity_damping, cooling, dissipation, vorticity;
  double seed;
} fp_settings;

/* FlameSettings (src/engine/flame.ts:5-11); ramps: smoke, heat, expansion. */
typedef struct fp_flame {
  double lifespan, smoke_rate, heat_rate, expansion_rate;
  int32_t knot_count[3];
  double knots[3][FP_MAX_FLAME_KNOTS][2]; /* [lifetime, response] */
} fp_flame;

/* FuelSettings (src/engine/fuel.ts:3-8). */
typedef struct fp_fuel {
  int32_t enabled;
  double ignition_heat, burn_rate;
} fp_fuel;

/* SimulationSource (src/engine/types.ts:246-268); sphere and box shapes. `id` stands in
 * for the reference's string id: a new id asks for pool slots at once. */
typedef struct fp_source {
  uint64_t id;
  int32_t shape;
  double position[3], size[3];
  double flame, heat_rate, smoke_rate, fuel_rate;
  double velocity[3];
  double velocity_response, outward_speed;
  int32_t has_variation;
  double variation_period, variation_strength;
  /* FP_SHAPE_MESH: the mesh's transformed surface samples in world space, (x, y, z, weight)
   * each, and its surface area in m^2 (MeshSurface.ts, handles.ts:160-183). Ignored for
   * other shapes. */
  const double *samples;
  int32_t sample_count;
  double surface_area;
} fp_source;

/* SimulationForce (src/engine/types.ts:284-293). `targeted` = sources given (even none). */
typedef struct fp_force {
  int32_t type;
  double vector[3], center[3];
  double strength, scale, lift, inward;
  int32_t targeted;
  int32_t source_count;
  const fp_source *sources;
} fp_force;

/* SimulationCollider (src/engine/types.ts:271-281), sphere or box. */
typedef struct fp_collider {
  int32_t shape;     /* FP_SHAPE_SPHERE, FP_SHAPE_BOX or FP_SHAPE_MESH */
  double position[3], size[3];
  double axes[3][3]; /* the box's axes in world space (rows); world axes by default */
  /* FP_SHAPE_MESH: the mesh's occupancy of a box of field cells (MeshColliderMask,
   * MeshCollider.ts voxelize): cells[0] * cells[1] * cells[2] words, nonzero inside, x fastest,
   * for the cells from lo; position, size and axes are then unused. */
  const uint32_t *occupancy;
  int32_t lo[3], cells[3];
} fp_collider;

/* SimulationFrame (src/engine/types.ts:295-302). */
typedef struct fp_frame {
  fp_settings simulation;
  fp_flame combustion;
  int32_t has_fuel;
  fp_fuel fuel;
  int32_t source_count;
  const fp_source *sources;
  int32_t force_count;
  const fp_force *forces;
  int32_t collider_count;
  const fp_collider *colliders;
} fp_frame;

/* The members FireSimulation.stats reads (FireSimulation.ts:295-334). */
typedef struct fp_stats {
  double time;
  int64_t steps;
  int64_t active_bricks;    /* activeBrickCount at the latest measurement */
  int64_t active_voxels;    /* active_bricks * brick_size^3 */
  int64_t pool_slots;       /* poolSlots */
  int64_t used_slots;       /* usedSlots */
  int32_t pool_limited;     /* poolLimited: gridLimited */
  int32_t tile_capacity;    /* tile slots the per-brick arrays hold */
  int32_t tile_count;       /* tiles in the directory */
  int32_t pool_side;        /* slots per row and rows per layer */
  int64_t memory_bytes;     /* device bytes this simulation holds (memoryBytes) */
  int64_t mesh_memory_bytes;  /* mesh emission records and brick tables (meshMemoryBytes,
                                 FluidSimulation.ts:849-851) */
  int64_t force_memory_bytes; /* force records, regions and mesh coverage (forceMemoryBytes,
                                 FluidSimulation.ts:846-848) */
} fp_stats;

/* The solver fields a debug view shows (FireSimulation.ts:125-127). */
#define FP_SOLVER_FIELD_NONE 0
#define FP_SOLVER_FIELD_PRESSURE 1
#define FP_SOLVER_FIELD_DIVERGENCE 2 /* the right-hand side h^2 (div u - expansion) solved for */

/* renderBricks (FluidSimulation.ts:163-189, :781-795): what the renderer marches. boxes:
 * 6 words each (first brick x, y, z; last brick x, y, z), one per cluster of tiles with
 * something to draw; tiles: 4 words per tile slot (tile x, y, z, the box of its cluster or
 * -1). The device tables stay on the simulation's device: pages (each brick id's page or
 * -1, brick_count words), the active bricks of the latest step (active_count ids) and the
 * bricks it zeroed (zeroed_count (id, payload) pairs). Buffers are caller-owned; NULL skips
 * one; *_count is what exists, written up to the matching capacity. */
typedef struct fp_render_bricks {
  int32_t version; /* changes whenever the boxes or the tiles do */
  int32_t *boxes;
  int32_t *tiles;
  int32_t box_capacity, tile_capacity;
  int32_t box_count, tile_count;
  const int32_t *pages_device;
  const int32_t *active_device;
  const int32_t *zeroed_device;
  int32_t brick_count, active_count, zeroed_count;
} fp_render_bricks;

/* Integer tables of the latest step, the bit-exact parity surface: every brick id's
 * page, entry and record, the free stack, pool state and the sorted active and zeroed
 * sets. Buffers are caller-owned; each *_count is the entries written. */
typedef struct fp_tables {
  int32_t *pages;       /* [brick_count] */
  int32_t *entries;     /* [brick_count] */
  int32_t *records;     /* [brick_count * 28] */
  uint32_t *free_slots; /* [pool_slots] */
  int32_t *active;      /* [brick_count] sorted brick ids */
  int32_t *zeroed;      /* [2 * brick_count*2] (id, payload) pairs, sorted by id then payload */
  int32_t brick_count, free_count, failures, active_count, zeroed_count;
  int32_t pool_slots;
} fp_tables;

/* The camera of the headless ray-march (scene-volume.wgsl fireProSceneVolume inputs).
 * inverse_view_projection maps clip coordinates to world space, column-major like
 * three.js Matrix4 (camera.matrixWorld x projectionMatrixInverse,
 * SceneVolumeRenderer.ts:452). Pixel (x, y) has uv ((x+.5)/width, (y+.5)/height), row 0
 * at the top. `position` is kept for callers that light a scene from the viewer. */
typedef struct fp_camera {
  double inverse_view_projection[16];
  double position[3];
  int32_t width, height;
} fp_camera;

/* RenderSettings (src/engine/types.ts) and RenderingOptions (src/library/options.ts),
 * with the scene's lights as SceneVolumeRenderer.readSceneLighting gathers them
 * (SceneVolumeRenderer.ts:298-325). Colors are linear RGB. */
typedef struct fp_render_settings {
  double fire_intensity, flame_opacity, smoke_density, scattering, shadow_density;
  double temperature, soot_glow;
  double smoke_color[3], fire_color[3];
  double sun_direction[3]; /* from the scene toward the directional light */
  double sun_intensity;    /* 0: no directional light, no shadows */
  double sun_color[3];     /* the directional light's color */
  double ambient[3];       /* sum of ambient light color x intensity */
  int32_t filter;          /* 0 trilinear, 1 quadratic, 2 cubic B-spline */
  int32_t ray_steps;       /* samples a ray takes at most across the boxes */
  int32_t lighting_divisor;/* 1, 2 or 4 field cells per lighting texel */
  int32_t half_resolution; /* 1: march at half resolution into depth slices first and resume
                              from them (VolumeSlices.ts; the reference's default) */
} fp_render_settings;

/* The render preparation of the latest render, for parity and inspection: the boxes
 * (8 words each: first brick xyz, first page-table entry, bricks per axis xyz, 0), the
 * page table, and for every listed brick (an active brick with a slot inside a box,
 * ascending id) its render field ((B+1)^3 float4 with apron), incident light
 * ((B/lighting_divisor+1)^3 float4 with apron) and block flags ((B/8)^3 bytes: bit 0
 * flame, bit 1 visible); and the 512-texel blackbody table. NULL buffers are skipped;
 * each *_count is what exists, written up to the matching capacity. */
typedef struct fp_render_tables {
  int32_t *boxes;
  int32_t *page_table;
  int32_t *bricks;
  float *field;
  float *light;
  uint8_t *occupancy;
  float *blackbody;
  int32_t box_capacity, page_capacity, brick_capacity;
  int32_t box_count, page_count, brick_count;
  int32_t light_work; /* 1 every listed brick lit, 2 only slots given this step, 0 none, -1 idle */
  /* smoke_divisor 2 or 4: each listed brick's smoothed coarse smoke, ((B/smoke_divisor)+1)^3
   * floats with apron (smoothCoarseSmoke); NULL skips it */
  float *smoke;
} fp_render_tables;

typedef struct fp_sim fp_sim;
typedef struct fp_batch fp_batch;

const char *fp_last_error(void);
const char *fp_version(void);

/* FluidSimulation.create (FluidSimulation.ts:359-475): pools, tiles, curl noise. */
int fp_sim_create(const fp_grid *grid, fp_sim **out);
/* FluidSimulation.step (FluidSimulation.ts:1367-1373): one substep of `dt` in (0, 1/30].
 * Returns once the step's activity readback has landed (FluidSimulation.ts:1514-1527); the
 * forces, pressure solve and projection may still run on the simulation's stream. Every
 * other entry point waits for them before it reads or changes the simulation. */
int fp_sim_step(fp_sim *sim, const fp_frame *frame, double dt);
/* FluidSimulation.reset (FluidSimulation.ts:768-779). */
int fp_sim_reset(fp_sim *sim);
/* FluidSimulation.dispose (FluidSimulation.ts:1780-1802); NULL is accepted. */
void fp_sim_destroy(fp_sim *sim);
/* The cutoff, voxelBudget and ground setters (FireSimulation.ts:457-459). */
int fp_sim_set_cutoff(fp_sim *sim, double cutoff);
int fp_sim_set_budget(fp_sim *sim, double max_voxels);
int fp_sim_set_ground(fp_sim *sim, int32_t ground);
/* The getters FireSimulation reads (FireSimulation.ts:296-333). */
int fp_sim_stats(const fp_sim *sim, fp_stats *out);
/* readBricks (FluidSimulation.ts:1749-1779): min corner of each slotted brick, meters,
 * in brick id order; `*count` in: capacity of `mins` (triples), out: bricks. */
int fp_sim_read_bricks(fp_sim *sim, double *mins, int32_t *count);
/* showSolverField (FluidSimulation.ts:1532-1559, called at FireSimulation.ts:211, :358):
 * with FP_SOLVER_FIELD_PRESSURE or _DIVERGENCE, copy that field of the latest step into the
 * simulation's solver-field view now and after every step (batched steps included);
 * FP_SOLVER_FIELD_NONE stops the copies. */
int fp_sim_show_solver_field(fp_sim *sim, int32_t field);
/* The solver-field view of the active bricks in active-list order, velocity-grid cells
 * (one float each) per brick; returns the brick count. FP_ESTATE when no field is shown. */
int fp_sim_read_solver_field(fp_sim *sim, int32_t *ids, float *values, int32_t capacity_bricks);
/* renderBricks (see fp_render_bricks). */
int fp_sim_render_bricks(fp_sim *sim, fp_render_bricks *out);
/* Integer tables of the latest step (see fp_tables). */
int fp_sim_read_tables(fp_sim *sim, fp_tables *out);
/* Cell values of the active bricks in active-list order: the field grid's four channels
 * (smoke, heat, lifetime, fuel) per cell, the velocity grid's three MAC components, the
 * pressure, and the expansion rate. NULL skips a field. Each buffer holds
 * active_count * cells-per-brick * channels floats; `ids` receives the brick ids. With
 * smoke_divisor 2 or 4 the fields' smoke channel holds the step's smoke production, and
 * fp_sim_read_smoke reads the density. */
int fp_sim_read_state(fp_sim *sim, int32_t *ids, float *fields, float *velocity,
                      float *pressure, float *expansion, int32_t capacity_bricks);
/* The coarse smoke grid's density (smoke_divisor 2 or 4; smoke.wgsl smokeDensity) of the
 * active bricks in active-list order, (brick_size / smoke_divisor)^3 floats per brick;
 * returns the brick count. FP_EINVAL for a full-resolution simulation. */
int fp_sim_read_smoke(fp_sim *sim, int32_t *ids, float *density, int32_t capacity_bricks);
/* Overwrite the cell values of the active bricks, same layout as fp_sim_read_state. */
int fp_sim_write_state(fp_sim *sim, const float *fields, const float *velocity);
/* The device stream the solver enqueues on (hipStream_t), for callers that time it or
 * wait on it (a step's tail is queued there when fp_sim_step returns). */
void *fp_sim_stream(fp_sim *sim);
/* Timing of the next steps, read back as averages in ms: enable 1 records device events
 * around every stage (and the render's) plus the host phases of a step (host_*); enable 2
 * records events around the dominant stage (evolve_fields) only, so the others run
 * without event gaps; 0 stops. Enabling clears earlier averages. */
int fp_sim_profile(fp_sim *sim, int32_t enable);
int fp_sim_profile_read(fp_sim *sim, const char **names, double *ms, int32_t *count);

/* Scene lights (SceneLights.sample, src/library/SceneLights.ts:127-218, and collectLights,
 * src/engine/shaders/scene-lights.wgsl:36-106): where the flames of the latest step light
 * the scene from. Every active brick samples its fields on a 4^3 lattice, weighs the
 * samples by the luminance of their emission (the render settings' flame and smoke
 * parameters) and adds them to the anchor nearest the brick's center, height above an
 * anchor counting a quarter (ties to the lower index). anchors: count world positions
 * (x, y, z). sums: count x 8 floats: the luminance, the luminance-weighted offset (x, y, z)
 * from the anchor, r, g, b, 0, summed in a fixed order. All zero before the first step or
 * without render boxes. */
int fp_sim_scene_lights(fp_sim *sim, const fp_render_settings *settings, const double *anchors, int32_t count,
                        float *sums);
/* fp_sim_scene_lights without allocating device memory: no counterpart in the reference.
 * For callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same results, bit for bit; its buffers are a fixed
 * 2.8 MiB device array per GPU, held until the process ends, so it takes at most 64
 * anchors over at most 65536 active bricks (FP_EINVAL beyond). Calls on one GPU share that
 * array: calls that overlap on one GPU race. */
int fp_sim_scene_lights_noalloc(fp_sim *sim, const fp_render_settings *settings, const double *anchors,
                                int32_t count, float *sums);

/* Batched stepping (the serving form; no counterpart in the reference, whose simulations
 * step one at a time). fp_batch_step advances `count` simulations, all of one brick size,
 * velocity divisor and pressure policy (each with its own tolerance and cycle budget), by
 * one step each with frames[i] for sims[i]: their host halves run in turn, then every
 * stage is one launch of persistent workgroups over all of their active bricks, defect
 * correction stops each simulation on the device when it has converged, and one readback
 * returns all their activity. The launch sequence is captured once as a HIP graph per
 * layout and replayed. Each simulation ends exactly as fp_sim_step would have left it. Any
 * set of simulations may be passed on each call (continuous batching); none may be
 * stepped or rendered elsewhere during the call. */
int fp_batch_create(fp_batch **out);
void fp_batch_destroy(fp_batch *batch);
int fp_batch_step(fp_batch *batch, fp_sim *const *sims, const fp_frame *frames, int32_t count, double dt);
/* Wall time of the latest fp_batch_step, host halves included, in ms. */
int fp_batch_last_ms(const fp_batch *batch, double *ms);
/* Its phases in ms: ms[0] host halves, [1] argument packing, [2] launches, [3] the wait for
 * the device, [4] the activity readbacks. */
int fp_batch_phases(const fp_batch *batch, double *ms);
/* 1 (the default) replays captured HIP graphs; 0 launches every kernel on each step. A batch
 * created under the rocprofv3 tool (ROCP_TOOL_LIBRARIES or a rocprofiler LD_PRELOAD set)
 * starts with 0. */
int fp_batch_use_graphs(fp_batch *batch, int32_t enable);

/* The pressure policy of the next steps. Defect correction keeps the solution x in FP32,
 * starts from x = 0 and repeats r = b - A x (FP32, each row's dot product and the norms in
 * FP64), e = one V-cycle on
 * A e = r (arithmetic FP32, levels stored at the mode's precision, coarsest level in
 * FP64), x += e, until |r| / |b| <= tolerance or max_cycles V-cycles have run. With
 * FP_PRESSURE_FP32 and max_cycles 1 it is FP_PRESSURE_VCYCLE but for the coarsest
 * level's FP64 sweeps. No counterpart in the reference, which runs one V-cycle. */
int fp_sim_set_pressure(fp_sim *sim, int32_t mode, double tolerance, int32_t max_cycles);
/* V-cycles the latest step ran and its final relative residual (-1 for FP_PRESSURE_VCYCLE). */
int fp_sim_pressure_report(const fp_sim *sim, int32_t *cycles, double *relative_residual);

/* The ray-march of the simulation's latest step (a12, a13): brings the render
 * preparation up to date on VolumeLighting.encode's schedule (VolumeLighting.ts:249-375)
 * and marches every pixel (scene-volume.wgsl:38-48, :146-167; volume-march.wgsl). Writes
 * width x height premultiplied (radiance rgb, 1 - transmittance) floats, row-major, to host
 * memory. Replaces SceneVolumeRenderer.prepare + the NodeMaterial's fragment pass. */
int fp_sim_render(fp_sim *sim, const fp_camera *camera, const fp_render_settings *settings, float *rgba);
/* The same into device memory (hipMalloc'ed, width*height*4 floats), enqueued on the
 * simulation's stream without waiting. */
int fp_sim_render_device(fp_sim *sim, const fp_camera *camera, const fp_render_settings *settings,
                         float *rgba_device);
int fp_sim_read_render(fp_sim *sim, fp_render_tables *out);
/* The latest render's sample counts, for measurement (no counterpart in the reference):
 * counts[0] the half-resolution slices' sample positions looked up in the occupancy and
 * counts[1] those reconstructed and shaded, counts[2] and counts[3] the same of the
 * full-resolution march. */
int fp_sim_render_stats(fp_sim *sim, uint64_t *counts);

#ifdef __cplusplus
}
#endif
#endif
