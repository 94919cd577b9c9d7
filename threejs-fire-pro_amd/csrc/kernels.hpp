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

// Multigrid level storage: brick levels per pool slot, tile levels per tile slot.
struct Level {
  float *p, *q, *rhs;
  uint8_t *topo, *eflag;
  int side, tile;
};
constexpr int MAX_LEVELS = 8;
struct MGArgs {
  Dir dir;
  Layout L;
  Params P;
  const int *counts;
  const int *active;
  const int *zeroList;
  const uint32_t *solids;
  Level lv[MAX_LEVELS];
  int bottom;
  float weight;  // axisWeights, equal on every axis for cubic cells
};

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
__global__ void k_mg_init(MGArgs m);
__global__ void k_mg_relax_bricks(MGArgs m, int level);
__global__ void k_mg_relax_tiles(MGArgs m, int level);
__global__ void k_mg_restrict_bricks(MGArgs m, int level);
__global__ void k_mg_restrict_tiles(MGArgs m, int level);
__global__ void k_mg_prolong_bricks(MGArgs m, int level);
__global__ void k_mg_prolong_tiles(MGArgs m, int level);
__global__ void k_mg_bottom(MGArgs m, int iterations);

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
