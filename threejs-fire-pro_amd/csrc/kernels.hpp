// Kernel argument blocks and launch entry points shared by the runtime and the kernels.
#pragma once
#include "sampling.hpp"

namespace fp {

// Most bricks (or tile slots) one multigrid workgroup takes at once.
constexpr int MG_GROUP = 32;

struct AllocArgs {
  Dir dir;
  Layout L;
  Params P;
  int brickCount, blocks;
  int *pages, *entries, *records;
  uint32_t *brickState, *activity;
  uint32_t *stateTmp, *rankTmp, *blockSums, *blockPrefix;
  uint32_t *freeSlots;
  int *poolState;  // free, failures
  int held;        // free slots the voxel budget holds back
  int *counts;     // active, fresh, free before, freed, requests, freed-zeroed
  int *active, *zeroList;
  const Emitter *emitters;
  const Collider *colliders;
  const uint32_t *colliderMasks;  // mesh colliders' occupancy, one bit per cell
  uint32_t *solids;
  int colliderCount;
  int buildSolids;  // draw the colliders into solid bits this step
  int solidsEvery;  // into every active slot's, else the fresh slots'
  const int *meshTables = nullptr;  // per brick id its mesh emission table or -1; null without mesh sources
  Pools p;
};

// Arguments of the per-cell kernels over the active bricks.
struct StepArgs {
  Dir dir;
  Layout L;
  Params P;
  Pools p;
  const int *counts;
  const int *active;
  uint32_t *activity;
  int brickCount;
  const Emitter *emitters;
  const ForceRec *forces;
  const Region *regions;
  // Mesh emission (fluid-sources.wgsl:4-22): per brick id a table index or -1, then per
  // table a record index per field cell; records are (flame, heat, smoke, fuel) and
  // (motion xyz, response), record 0 all zero. Null without mesh sources.
  const int *meshTables = nullptr;
  const float4 *meshRecords = nullptr;
  // Velocity cells that forces' mesh targets cover (force-fields.wgsl:12-47): per brick id
  // the first word of its entries or 0xffffffff, then a count and per entry a force index
  // and one bit per velocity cell. Null without such targets.
  const uint32_t *forceCoverage = nullptr;
  const float *pressure = nullptr;    // the solved pressure the projection reads
  const float *pressureLo = nullptr;  // its low part when it is a pair hi + lo (defect correction), else null
};

// Storage of the multigrid levels' iterates and right-hand sides. Arithmetic is FP32 in
// every case; values are rounded to nearest even when stored.
struct StoreF32 {
  using T = float;
  static __device__ __forceinline__ float ld(float v) { return v; }
  static __device__ __forceinline__ float st(float v) { return v; }
};
struct StoreBF16 {
  using T = uint16_t;  // bfloat16 bits
  static __device__ __forceinline__ float ld(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
  static __device__ __forceinline__ uint16_t st(float v) {
    uint32_t b = __float_as_uint(v);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
  }
};
struct StoreF16 {
  using T = _Float16;
  static __device__ __forceinline__ float ld(_Float16 v) { return (float)v; }
  static __device__ __forceinline__ _Float16 st(float v) { return (_Float16)v; }
};

// Multigrid level storage: brick levels per pool slot, tile levels per tile slot.
template <class T> struct LevelT {
  T *p, *q, *rhs;
  uint8_t *topo, *eflag;
  int side, tile;
};
constexpr int MAX_LEVELS = 8;
template <class T> struct MGArgsT {
  Dir dir;
  Layout L;
  Params P;
  const int *counts;
  const int *active;
  const int *zeroList;
  const uint32_t *solids;
  LevelT<T> lv[MAX_LEVELS];
  int bottom;
  float weight;     // axisWeights, equal on every axis for cubic cells
  int topologyAll;  // the geometry changed: rebuild every cell's topology
  int topologyTiles = 0;  // the tile set changed: rebuild the tile levels' topology
  // The levels below the first tile level run in k_mg_coarse_cycle (at most TC_MAX tile
  // slots): their per-level kernels and the bottom kernels skip this simulation.
  int tilesFused = 0;
  // Defect correction: right-hand side b, the solution x as an unevaluated sum of two FP32
  // values dcX + dcXlo, residual r, per-brick partial sums and the norms (sum r^2, sum b^2,
  // max |r|, scale, relative residual, finished, cycles). Once `finished` is set every
  // later kernel of the solve skips this simulation.
  const float *dcB = nullptr;
  float *dcX = nullptr, *dcXlo = nullptr, *dcR = nullptr;
  double *dcPartials = nullptr, *dcNorms = nullptr;
  double dcTol = 0.0;   // finish when the relative residual is at most this
  int dcMaxCycles = 0;  // or after this many V-cycles
  int dcSlots = 0;      // pool slots of x
};
template <class T> __device__ __forceinline__ bool dcFinished(const MGArgsT<T> &m) {
  return m.dcNorms != nullptr && m.dcNorms[5] != 0.0;
}
using Level = LevelT<float>;
using MGArgs = MGArgsT<float>;

// Ragged batches. A batched launch (k_*_batch) advances many simulations: its workgroups
// stride over virtual blocks, starts[i] being simulation i's first and starts[n] the total,
// each a multiple of 8 so that the XCD-aware mapping holds inside every simulation; each
// virtual block runs the single-simulation kernel's body with that simulation's arguments,
// read from device memory. The grid is fixed (persistent workgroups), so one captured
// launch serves any mix of simulations.
__device__ __forceinline__ int batchOwner(const int *starts, int n, int v) {
  int lo = 0, hi = n - 1;  // the last i with starts[i] <= v
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
#define FP_BATCH_LOOP(NAME, CALL)                                  \
  const int total = starts[n];                                     \
  for (int v = blockIdx.x; v < total; v += gridDim.x) {            \
    const int sim = __builtin_amdgcn_readfirstlane(batchOwner(starts, n, v)); \
    CALL;                                                          \
    __syncthreads();                                               \
  }
// A kernel body NAME_body(args, virtual block[, one more parameter]) as the single-simulation
// kernel k_NAME and the batched k_NAME_batch.
#define FP_KERNEL_PAIR(NAME, BOUNDS, ARGS)                                                                  \
  __global__ void __launch_bounds__(BOUNDS) k_##NAME(ARGS a) { NAME##_body(a, (int)blockIdx.x); }           \
  __global__ void __launch_bounds__(BOUNDS) k_##NAME##_batch(const ARGS *args, const int *starts, int n) {  \
    FP_BATCH_LOOP(NAME, NAME##_body(args[sim], v - starts[sim]))                                             \
  }
#define FP_KERNEL_PAIR1(NAME, BOUNDS, ARGS, PTYPE, PNAME)                                                     \
  __global__ void __launch_bounds__(BOUNDS) k_##NAME(ARGS a, PTYPE PNAME) { NAME##_body(a, (int)blockIdx.x, PNAME); } \
  __global__ void __launch_bounds__(BOUNDS) k_##NAME##_batch(const ARGS *args, const int *starts, int n, PTYPE PNAME) { \
    FP_BATCH_LOOP(NAME, NAME##_body(args[sim], v - starts[sim], PNAME))                                      \
  }
// As FP_KERNEL_PAIR1, compiled for WAVES resident wavefronts a SIMD (at most 512 / WAVES
// VGPRs a lane).
#define FP_KERNEL_PAIR1_OCC(NAME, BOUNDS, WAVES, ARGS, PTYPE, PNAME)                                          \
  __global__ void __launch_bounds__(BOUNDS) __attribute__((amdgpu_waves_per_eu(WAVES)))                     \
      k_##NAME(ARGS a, PTYPE PNAME) { NAME##_body(a, (int)blockIdx.x, PNAME); }                               \
  __global__ void __launch_bounds__(BOUNDS) __attribute__((amdgpu_waves_per_eu(WAVES)))                     \
      k_##NAME##_batch(const ARGS *args, const int *starts, int n, PTYPE PNAME) {                              \
    FP_BATCH_LOOP(NAME, NAME##_body(args[sim], v - starts[sim], PNAME))                                      \
  }
#define FP_DECLARE_PAIR(NAME, ARGS) \
  __global__ void k_##NAME(ARGS a);   \
  __global__ void k_##NAME##_batch(const ARGS *args, const int *starts, int n);
#define FP_DECLARE_PAIR1(NAME, ARGS, PTYPE)  \
  __global__ void k_##NAME(ARGS a, PTYPE); \
  __global__ void k_##NAME##_batch(const ARGS *args, const int *starts, int n, PTYPE);

FP_DECLARE_PAIR(build_bricks, AllocArgs)
FP_DECLARE_PAIR(scan_blocks, AllocArgs)
FP_DECLARE_PAIR(free_bricks, AllocArgs)
FP_DECLARE_PAIR(allocate_bricks, AllocArgs)
FP_DECLARE_PAIR(link_bricks, AllocArgs)
FP_DECLARE_PAIR(build_solids, AllocArgs)
__global__ void k_link_fresh(AllocArgs a, int linkBlocks, int chunks, int solids);
FP_DECLARE_PAIR1(zero_fresh, AllocArgs, int)

FP_DECLARE_PAIR(advect_velocity, StepArgs)
FP_DECLARE_PAIR(correct_velocity, StepArgs)
FP_DECLARE_PAIR1(advect_scalars, StepArgs, int)
FP_DECLARE_PAIR1(evolve_fields, StepArgs, int)
FP_DECLARE_PAIR(advect_smoke, StepArgs)
FP_DECLARE_PAIR(evolve_smoke, StepArgs)
FP_DECLARE_PAIR(curl, StepArgs)
FP_DECLARE_PAIR(cell_forces, StepArgs)
FP_DECLARE_PAIR(forces, StepArgs)
FP_DECLARE_PAIR(divergence, StepArgs)
FP_DECLARE_PAIR(project, StepArgs)

// Multigrid kernels. A level's iterate alternates between its p and q arrays; bit L of
// `flips` says that level L's current iterate is in q. Level kernels take
// `code` = level | flips << 8.
FP_DECLARE_PAIR1(topology_bricks, MGArgs, int)
FP_DECLARE_PAIR1(topology_tiles, MGArgs, int)
FP_DECLARE_PAIR(eflags_bricks, MGArgs)
__global__ void k_fresh_topology(MGArgs m);
__global__ void k_near_topology(MGArgs m, const int *tiles, int count);
__global__ void k_topology_tiles_list(MGArgs m, const int *tiles, int count);
__global__ void k_moved_topology(MGArgs m, int fresh, const int *tiles, int nearCount, int tileCount);
FP_DECLARE_PAIR(eflags_tiles, MGArgs)
__global__ void k_mg_flags_init(MGArgs m, int brickGroups);  // eflags of every level + mg_init, one launch
#define FP_DECLARE_MG(NAME)                                                                       \
  template <class St> __global__ void k_##NAME(MGArgsT<typename St::T> a, int code);              \
  template <class St>                                                                             \
  __global__ void k_##NAME##_batch(const MGArgsT<typename St::T> *args, const int *starts, int n, int code);
FP_DECLARE_MG(mg_init)
FP_DECLARE_MG(mg_relax_bricks)
FP_DECLARE_MG(mg_relax_tiles)
FP_DECLARE_MG(mg_restrict_bricks)
FP_DECLARE_MG(mg_restrict_tiles)
FP_DECLARE_MG(mg_prolong_bricks)
FP_DECLARE_MG(mg_prolong_tiles)
FP_DECLARE_MG(mg_bottom)    // relaxCoarse: 19 sweeps in one workgroup, up to BOTTOM64_TILES tile slots
FP_DECLARE_MG(mg_bottom64)  // the same in FP64, up to BOTTOM64_TILES tile slots
// The first tile level's two sweeps and the restriction below it, and the prolongation into
// it and its three sweeps, one workgroup per tile slot over the tile and a three-cell halo.
FP_DECLARE_MG(mg_tiles_down)
FP_DECLARE_MG(mg_tiles_up)
FP_DECLARE_MG(mg_bricks_down)
FP_DECLARE_MG(mg_bricks_up)
// The V-cycle below the first tile level in one workgroup, for simulations of at most TC_MAX
// tile slots (MGArgsT::tilesFused).
FP_DECLARE_MG(mg_coarse_cycle)
// Active bricks a k_link_bricks workgroup of 256 lanes links (32 lanes each).
constexpr int LINK_BRICKS = 8;
constexpr int TC_MAX = 512;
constexpr int TC_CELLS = TC_MAX * 9;  // 2^3 + 1 cells per tile slot
constexpr int TC_LANES = 1024;
// Defect correction: r = b - A x (FP32 with FP64 row dot products) and per-brick FP64 sums;
// their totals, the power-of-two scale and the stopping test of cycle `code` (one
// workgroup per simulation); level 0's right-hand side r / scale; x += scale * (level 0's
// solution).
FP_DECLARE_MG(dc_start)  // x = 0 and the finished flag cleared, before the loop
FP_DECLARE_MG(dc_residual)
FP_DECLARE_MG(dc_norms)
FP_DECLARE_MG(dc_scale)
FP_DECLARE_MG(dc_update)
// The batched solve's per-level virtual-block prefixes over unfinished members.
template <class St>
__global__ void k_level_prefix_batch(const MGArgsT<typename St::T> *args, int n, int *out);
constexpr int BOTTOM64_TILES = 8192;
constexpr int BOTTOM_ITERATIONS = 19;  // PressureSolver.ts:326

// The render's parameters (volume-common.wgsl RenderParams, packed at
// SceneVolumeRenderer.ts:329-345, and the camera uniforms of :440-456).
struct RenderParams {
  float voxel, brickSize, texel;
  int shift, lightShift, lightDivisor;
  float flameOpacity;
  float optical[4];  // smoke density, fire intensity, shadow density, anisotropy
  float light[4];    // direction toward the light, intensity
  float smoke[3], sootGlow;
  float fire[4];     // tint, temperature
  float sceneLight[3], ambient[3];
  float filter, raySteps, time;
  float ivp[16];     // inverse view-projection, column-major
  int width, height;
};

// The render preparation's tables and pools. Render field and light slots carry a
// one-texel apron at their far faces: (B+1)^3 and (B/ld+1)^3 float4 per slot.
struct RenderArgs {
  int B, sF, side, nF;
  int E, nR;                // render field slot edge and texels
  int lside, LE, nL;        // light brick edge, slot edge, texels
  int perSide, nO;          // occupancy blocks per axis, per slot
  const int4 *boxes;        // two words each: first brick + first entry, size
  int boxCount;
  int *table;               // page tables
  const int4 *tiles;        // each tile slot's coordinates and box, or -1
  const int *pages;         // each brick id's page
  const int *counts;        // [0] active, [1] fresh
  const int *active;
  const int *zeroList;
  int brickCount;
  const float4 *F;
  float4 *R, *Lt;
  uint8_t *O;
  const float4 *lut;
  RenderParams P;
  // Coarse smoke (smokeDivisor 2 or 4): the smoke brick's shift, edge and cells, the render
  // smoke's slot edge (edge + 1) and texels, the simulation's density and its smoothed copy
  // with one-texel aprons, and the directory (the scene lights read neighbors through it).
  int coarse;
  int sS, bs, nS, ES, nRS;
  const float *S;
  float *RS;
  Dir dir;
};

__global__ void k_export_pages(RenderArgs r);
__global__ void k_smooth_smoke(RenderArgs r, int chunks, int staged);
__global__ void k_smooth_coarse(RenderArgs r);
__global__ void k_occupancy(RenderArgs r);
__global__ void k_light(RenderArgs r, int fresh);
// The march and the half-resolution slices add their samples to stats (may be null): [0, 1]
// the slices' positions visited and samples shaded, [2, 3] the march's.
__global__ void k_march(RenderArgs r, float4 *out, const float4 *slices, int sw, int sh, unsigned long long *stats);
__global__ void k_slices(RenderArgs r, float4 *slices, int sw, int sh, const int *order, uint32_t *cost,
                         unsigned long long *stats);
__global__ void k_order_tiles(const uint32_t *cost, int tiles, int *order);
// Scene lights: per-brick totals by nearest anchor, then per-anchor totals in a fixed order
// (blocks of LIGHT_BLOCK bricks).
constexpr int LIGHT_BLOCK = 256;
__global__ void k_light_bricks(RenderArgs r, const int *tileRec, const float4 *anchors, int count, float *partials,
                               int *brickAnchor);
__global__ void k_light_blocks(const int *counts, const float *partials, const int *brickAnchor, int count,
                               float *blockSums);
__global__ void k_light_total(const int *counts, const float *blockSums, int count, float *sums);

// Lane i of the activity readback (FluidSimulation.ts:1514-1525): the counts, the pool state
// and the tiles' content flags; the lane that reads the failure counter restarts it.
__device__ __forceinline__ void readbackLane(const int *counts, int *poolState, const uint32_t *tileFlags, int cap,
                                             int *out, int i) {
  if (i == 0) {
    out[0] = counts[0];
    out[1] = poolState[0];
    out[2] = poolState[1];
    out[3] = counts[1] + counts[5];
    poolState[1] = 0;
  }
  if (i < cap) out[4 + i] = (int)tileFlags[i];
}
// The curl (workgroups below curlBlocks) and the readback (the rest) in one launch.
__global__ void k_curl_readback(StepArgs s, int curlBlocks, const int *counts, int *poolState,
                                const uint32_t *tileFlags, int cap, int *out);

// Loads the home brick's record into LDS; every lane of the block calls it.
__device__ __forceinline__ void loadHome(const Dir &dir, int id, int *rec, Home &h, int side) {
  if (threadIdx.x < 27) {
    int e = dir.records[(size_t)id * RECORD_WORDS + threadIdx.x];
    rec[threadIdx.x] = e;
    rec[27 + homeSlotWord(threadIdx.x)] = entrySlot(e, side);
  }
  const int *tr = dir.tileRec + (size_t)(id >> 6) * TILE_RECORD_WORDS;
  int l = id & 63;
  h.bx = tr[0] * 4 + (l & 3);
  h.by = tr[1] * 4 + ((l >> 2) & 3);
  h.bz = tr[2] * 4 + (l >> 4);
  h.rec = rec;
  h.slot = rec + 27;
  __syncthreads();
}

}  // namespace fp
