// Kernel argument blocks and launch entry points shared by the runtime and the kernels.
#pragma once
#include "sampling.hpp"

namespace fp {

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
  uint32_t *solids;
  int colliderCount;
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
  float weight;  // axisWeights, equal on every axis for cubic cells
};
using Level = LevelT<float>;
using MGArgs = MGArgsT<float>;

__global__ void k_build_bricks(AllocArgs a);
__global__ void k_scan_blocks(AllocArgs a);
__global__ void k_free_bricks(AllocArgs a);
__global__ void k_allocate_bricks(AllocArgs a);
__global__ void k_link_bricks(AllocArgs a);
__global__ void k_build_solids(AllocArgs a, int everySlot);
__global__ void k_zero_fresh(AllocArgs a, Pools p, int chunks);

__global__ void k_advect_velocity(StepArgs s);
__global__ void k_correct_velocity(StepArgs s);
__global__ void k_advect_scalars(StepArgs s, int chunks);
__global__ void k_evolve_fields(StepArgs s, int chunks);
__global__ void k_curl(StepArgs s);
__global__ void k_cell_forces(StepArgs s, float *out);
__global__ void k_forces(StepArgs s, const float *cellF);
__global__ void k_divergence(StepArgs s, float *rhs0);
__global__ void k_project(StepArgs s, const float *pressure);

__global__ void k_topology_bricks(MGArgs m, int fromZeroList);
__global__ void k_topology_tiles(MGArgs m, int level);
__global__ void k_eflags_bricks(MGArgs m);
__global__ void k_eflags_tiles(MGArgs m, int level);
template <class St> __global__ void k_mg_init(MGArgsT<typename St::T> m);
template <class St> __global__ void k_mg_relax_bricks(MGArgsT<typename St::T> m, int level);
template <class St> __global__ void k_mg_relax_tiles(MGArgsT<typename St::T> m, int level);
template <class St> __global__ void k_mg_restrict_bricks(MGArgsT<typename St::T> m, int level);
template <class St> __global__ void k_mg_restrict_tiles(MGArgsT<typename St::T> m, int level);
template <class St> __global__ void k_mg_prolong_bricks(MGArgsT<typename St::T> m, int level);
template <class St> __global__ void k_mg_prolong_tiles(MGArgsT<typename St::T> m, int level);
template <class St> __global__ void k_mg_bottom(MGArgsT<typename St::T> m, int iterations);
// The coarsest level in FP64 (any tile count up to BOTTOM64_TILES), stored back rounded.
constexpr int BOTTOM64_TILES = 8192;
template <class St> __global__ void k_mg_bottom64(MGArgsT<typename St::T> m, int iterations);
// Defect correction: r = b - A x (FP32) with per-brick FP64 sums of r^2, b^2 and the
// largest |r|; their totals and a power-of-two scale; level 0's right-hand side r / scale;
// x += scale * (level 0's solution).
template <class St> __global__ void k_dc_residual(MGArgsT<typename St::T> m, const float *b, const float *x, float *r,
                                                  double *partials);
__global__ void k_dc_norms(const double *partials, const int *counts, double *out);
template <class St> __global__ void k_dc_scale(MGArgsT<typename St::T> m, const float *r, const double *norms);
template <class St> __global__ void k_dc_update(MGArgsT<typename St::T> m, float *x, const double *norms);

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
};

__global__ void k_export_pages(RenderArgs r);
__global__ void k_smooth_smoke(RenderArgs r, int chunks, int staged);
__global__ void k_occupancy(RenderArgs r);
__global__ void k_light(RenderArgs r, int fresh);
__global__ void k_march(RenderArgs r, float4 *out);

// Loads the home brick's record into LDS; every lane of the block calls it.
__device__ __forceinline__ void loadHome(const Dir &dir, int id, int *rec, Home &h, int side) {
  if (threadIdx.x < 27) {
    int e = dir.records[(size_t)id * RECORD_WORDS + threadIdx.x];
    rec[threadIdx.x] = e;
    rec[27 + threadIdx.x] = entrySlot(e, side);
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
