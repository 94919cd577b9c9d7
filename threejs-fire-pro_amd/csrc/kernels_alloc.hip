// Brick activation, deterministic slot allocation and stream compaction
// (fluid-allocation.wgsl:129-387). Slots are freed and granted in brick-id order: a block
// scan ranks each brick's free, request and keep within its 256-brick block, one workgroup
// scans the block totals, and the active list is compacted in brick-id order from those
// ranks, so every table is reproducible run to run.
#include "kernels.hpp"

namespace fp {

constexpr int RANK_BLOCK = 256;

// Whether an emitter's footprint, padded by the stencil and a step of motion, reaches
// brick b (sourceNear, fluid-allocation.wgsl:305-323).
__device__ __forceinline__ bool sourceNear(const AllocArgs &a, const int b[3]) {
  float size = (float)(1 << a.L.sF) * a.P.hf;
  float lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    lo[k] = (float)b[k] * size;
    hi[k] = ((float)b[k] + 1.0f) * size;
  }
  for (int i = 0; i < a.P.emitterCount; i++) {
    const Emitter &e = a.emitters[i];
    bool in = true;
    for (int k = 0; k < 3; k++) {
      float motion = fabsv(e.vel[k]) + fabsv(e.outward);
      float pad = fminv(size, 2.0f * a.P.hv + motion * a.P.dt);
      float extent = fmaxv(e.size[k], 0.03f) + pad;
      if (!(e.pos[k] + extent > lo[k] && e.pos[k] - extent < hi[k])) in = false;
    }
    if (in) return true;
  }
  return false;
}

// Per-lane ranks of three 0/1 flags within a 256-lane block, and the block totals.
__device__ __forceinline__ void blockRanks3(const bool f[3], uint32_t rank[3], uint32_t total[3], uint32_t *lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t inwave[3];
  for (int k = 0; k < 3; k++) {
    uint64_t m = __ballot(f[k]);
    inwave[k] = (uint32_t)__popcll(m & below);
    if (lane == 0) lds[k * 4 + wave] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  for (int k = 0; k < 3; k++) {
    uint32_t before = 0, sum = 0;
    for (int w = 0; w < 4; w++) {
      uint32_t v = lds[k * 4 + w];
      if (w < wave) before += v;
      sum += v;
    }
    rank[k] = before + inwave[k];
    total[k] = sum;
  }
}

// buildBricks (:129-175): each brick's state; per-block totals of frees, requests, keeps.
__device__ __forceinline__ void build_bricks_body(const AllocArgs &a, const int vb) {
  __shared__ uint32_t lds[12];
  const int index = vb * RANK_BLOCK + threadIdx.x;
  const int n = a.brickCount;
  if (index < a.dir.capacity) a.activity[n + index] = 0u;  // this step's tile content
  uint32_t state = 0;
  bool keep = false;
  if (index < n) {
    const int tile = index >> 6;
    const int *tr = a.dir.tileRec + (size_t)tile * TILE_RECORD_WORDS;
    if (tr[3] != 0) {
      int l = index & 63;
      int b[3] = {tr[0] * 4 + (l & 3), tr[1] * 4 + ((l >> 2) & 3), tr[2] * 4 + (l >> 4)};
      bool open = tr[3] != 2;
      bool listed = false;
      if (open) {
        // Whether any neighbor's content support reaches this brick (or a mesh source covers
        // one): the 27 reads issued together and OR-ed, the same answer as a scan that stops
        // at the first.
        uint32_t any = 0u;
#pragma unroll
        for (int k = 0; k < 27; k++) {
          const int x = k % 3 - 1, y = (k / 3) % 3 - 1, z = k / 9 - 1;
          const int nb[3] = {b[0] + x, b[1] + y, b[2] + z};
          const int dt[3] = {(nb[0] >> 2) - tr[0], (nb[1] >> 2) - tr[1], (nb[2] >> 2) - tr[2]};
          const int ntile = tr[4 + (dt[2] + 1) * 9 + (dt[1] + 1) * 3 + (dt[0] + 1)];
          if (ntile < 0) continue;
          const int nid = ntile * 64 + (((nb[2] & 3) * 4 + (nb[1] & 3)) * 4 + (nb[0] & 3));
          const int toward = (-z + 1) * 9 + (-y + 1) * 3 + (-x + 1);
          any |= (a.activity[nid] >> 3) & (1u << toward);
          if (a.meshTables && a.meshTables[nid] >= 0) any |= 1u;
        }
        listed = any != 0u;
      }
      if (open && !listed) listed = sourceNear(a, b);
      int page = a.pages[index];
      bool was = (a.brickState[index] & 1u) != 0u;
      state = (uint32_t)listed | ((uint32_t)was << 2);
      if (!listed && page >= 0) state |= BRICK_FREES;
      if (listed && page < 0) state |= BRICK_REQUESTS;
      keep = listed && page >= 0;
    }
  }
  bool f[3] = {(state & BRICK_FREES) != 0u, (state & BRICK_REQUESTS) != 0u, keep};
  uint32_t rank[3], total[3];
  blockRanks3(f, rank, total, lds);
  if (threadIdx.x == 0) {
    a.blockSums[vb * 3 + 0] = total[0];
    a.blockSums[vb * 3 + 1] = total[1];
    a.blockSums[vb * 3 + 2] = total[2];
  }
  if (index < n) {
    a.stateTmp[index] = state | ((uint32_t)keep << 5);
    a.rankTmp[index] = rank[0] | (rank[1] << 10) | (rank[2] << 20);
  }
}
FP_KERNEL_PAIR(build_bricks, 256, AllocArgs)


// One workgroup: exclusive prefix of the block totals, then the step's commit
// (finishBricks, :352-387): granted = min(requests, max(free + freed - held, 0)).
__device__ __forceinline__ void scan_blocks_body(const AllocArgs &a, const int vb) {
  __shared__ uint32_t part[3][1024];
  __shared__ uint32_t carry[3];
  const int blocks = a.blocks;
  if (threadIdx.x < 3) carry[threadIdx.x] = 0;
  __syncthreads();
  for (int base = 0; base < blocks; base += 1024) {
    int i = base + threadIdx.x;
    for (int k = 0; k < 3; k++) part[k][threadIdx.x] = i < blocks ? a.blockSums[i * 3 + k] : 0u;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      uint32_t v[3];
      for (int k = 0; k < 3; k++) v[k] = threadIdx.x >= (unsigned)off ? part[k][threadIdx.x - off] : 0u;
      __syncthreads();
      for (int k = 0; k < 3; k++) part[k][threadIdx.x] += v[k];
      __syncthreads();
    }
    if (i < blocks)
      for (int k = 0; k < 3; k++) a.blockPrefix[i * 3 + k] = carry[k] + part[k][threadIdx.x] - a.blockSums[i * 3 + k];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < 3; k++) carry[k] += part[k][1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int freed = (int)carry[0], requests = (int)carry[1], keeps = (int)carry[2];
    int free0 = a.poolState[0], held = a.held;
    int freeAfter = free0 + freed;
    int room = freeAfter - held > 0 ? freeAfter - held : 0;
    int granted = requests < room ? requests : room;
    a.poolState[0] = freeAfter - granted;
    a.poolState[1] += requests - granted;
    a.counts[0] = keeps + granted;  // active bricks
    a.counts[1] = granted;          // fresh slots
    a.counts[2] = free0;
    a.counts[3] = freed;
    a.counts[4] = requests;
    a.counts[5] = 0;                // freed slots appended to the zeroed list
  }
}
FP_KERNEL_PAIR(scan_blocks, 1024, AllocArgs)


// freeBricks (:179-195): freed slots go onto the stack in brick order.
__device__ __forceinline__ void free_bricks_body(const AllocArgs &a, const int vb) {
  const int index = vb * RANK_BLOCK + threadIdx.x;
  if (index >= a.brickCount) return;
  uint32_t state = a.stateTmp[index];
  if (!(state & BRICK_FREES)) return;
  int rank = (int)(a.blockPrefix[vb * 3] + (a.rankTmp[index] & 1023u));
  int page = a.pages[index];
  a.freeSlots[a.counts[2] + rank] = (uint32_t)pageToSlot(page, a.L.side);
  if (state & 4u) {
    int z = a.counts[1] + atomicAdd(&a.counts[5], 1);
    a.zeroList[2 * z] = index;
    a.zeroList[2 * z + 1] = page;
  }
  a.pages[index] = -1;
}
FP_KERNEL_PAIR(free_bricks, 256, AllocArgs)


// allocateBricks (:200-230): grants from the stack top in brick order, entries, the
// compacted active list and the activity reset.
__device__ __forceinline__ void allocate_bricks_body(const AllocArgs &a, const int vb) {
  const int index = vb * RANK_BLOCK + threadIdx.x;
  if (index >= a.brickCount) return;
  uint32_t state = a.stateTmp[index];
  uint32_t ranks = a.rankTmp[index];
  const uint32_t *pre = a.blockPrefix + vb * 3;
  const int granted = a.counts[1];
  int reqBefore = (int)(pre[1] + ((ranks >> 10) & 1023u));
  int keepBefore = (int)(pre[2] + ((ranks >> 20) & 1023u));
  int page = a.pages[index];
  bool fresh = false;
  if ((state & BRICK_REQUESTS) && reqBefore < granted) {
    int top = a.counts[2] + a.counts[3] - reqBefore;
    page = slotToPage((int)a.freeSlots[top - 1], a.L.side);
    a.pages[index] = page;
    a.zeroList[2 * reqBefore] = index;
    a.zeroList[2 * reqBefore + 1] = (int)((uint32_t)page | FRESH_SLOT);
    fresh = true;
  }
  if (page >= 0) a.active[keepBefore + (reqBefore < granted ? reqBefore : granted)] = index;
  a.brickState[index] = page >= 0 ? 1u : 0u;
  a.entries[index] = page >= 0 ? (page | (fresh ? FRESH_ENTRY : 0)) : -1;
  uint32_t flags = a.activity[index];
  a.activity[index] = (flags & EXPANDS) ? MAY_EXPAND : 0u;
}
FP_KERNEL_PAIR(allocate_bricks, 256, AllocArgs)


// linkBricks (:234-268): each active brick's 27 neighbor entries and the pressure flag.
// 32 lanes a brick, LINK_BRICKS bricks a 256-lane workgroup: lane k < 27 resolves neighbor
// k's entry, lane 27 the pressure flag from the eight below once they are in LDS; each
// brick's record is written as one run of 28 words.
__device__ __forceinline__ void link_bricks_body(const AllocArgs &a, const int vb) {
  __shared__ int ws[LINK_BRICKS][28];
  const int j = (int)threadIdx.x >> 5, k = (int)threadIdx.x & 31;
  const int i = vb * LINK_BRICKS + j;
  const bool live = i < a.counts[0];
  int index = 0, b[3] = {0, 0, 0};
  if (live && k < 28) {
    index = a.active[i];
    const int *tr = a.dir.tileRec + (size_t)(index >> 6) * TILE_RECORD_WORDS;
    const int l = index & 63;
    b[0] = tr[0] * 4 + (l & 3);
    b[1] = tr[1] * 4 + ((l >> 2) & 3);
    b[2] = tr[2] * 4 + (l >> 4);
    if (k < 27) {
      const int nb[3] = {b[0] + k % 3 - 1, b[1] + (k / 3) % 3 - 1, b[2] + k / 9 - 1};
      const int dt[3] = {(nb[0] >> 2) - tr[0], (nb[1] >> 2) - tr[1], (nb[2] >> 2) - tr[2]};
      const int ntile = tr[4 + (dt[2] + 1) * 9 + (dt[1] + 1) * 3 + (dt[0] + 1)];
      const int w = ntile < 0 ? -1 : a.entries[ntile * 64 + (((nb[2] & 3) * 4 + (nb[1] & 3)) * 4 + (nb[0] & 3))];
      ws[j][k] = w;
      a.records[(size_t)index * RECORD_WORDS + k] = w;
    }
  }
  __syncthreads();
  if (live && k == PRESSURE_FLAG) {
    int unknowns = 1;
    for (int c = 0; c < 8; c++) {
      const int dx = -(c & 1), dy = -((c >> 1) & 1), dz = -((c >> 2) & 1);
      if (a.P.ground && b[1] + dy < 0) continue;
      if (ws[j][(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] < 0) unknowns = 0;
    }
    a.records[(size_t)index * RECORD_WORDS + PRESSURE_FLAG] = unknowns;
  }
}
FP_KERNEL_PAIR(link_bricks, 256, AllocArgs)


// Colliders into the solid bits of one brick (buildSolids, :292-341): a workgroup per
// brick, a word of 32 cells per lane and pass.
__device__ __forceinline__ bool insideCollider(const Collider &c, const uint32_t *masks, float hf, int x, int y, int z) {
  if (c.box == 2) {
    // A mesh's occupancy of its box of cells (lighting.wgsl insideCollider, fluid-allocation.wgsl:274-282).
    const int q[3] = {x - c.lo[0], y - c.lo[1], z - c.lo[2]};
    if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= c.size[0] || q[1] >= c.size[1] || q[2] >= c.size[2]) return false;
    const uint32_t index = (uint32_t)((q[2] * c.size[1] + q[1]) * c.size[0] + q[0]);
    return ((masks[c.maskOffset + (index >> 5)] >> (index & 31u)) & 1u) != 0u;
  }
  float w[3] = {((float)x + 0.5f) * hf, ((float)y + 0.5f) * hf, ((float)z + 0.5f) * hf};
  float rel[3] = {w[0] - c.center[0], w[1] - c.center[1], w[2] - c.center[2]};
  float dd[3];
  for (int k = 0; k < 3; k++) {
    float proj = c.axes[k][0] * rel[0] + c.axes[k][1] * rel[1] + c.axes[k][2] * rel[2];
    dd[k] = fabsv(proj / c.extent[k]);
  }
  if (c.box == 1) return dd[0] < 1.0f && dd[1] < 1.0f && dd[2] < 1.0f;
  return dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] < 1.0f;
}

// One brick (its id and page) with a wavefront (`lane` its lane): a lane per cell and pass,
// the pass's 64 bits from a ballot (two words, cells in index order).
__device__ __forceinline__ void solidBrick(const AllocArgs &a, int index, int page, int lane) {
  const int *tr = a.dir.tileRec + (size_t)(index >> 6) * TILE_RECORD_WORDS;
  const int l = index & 63;
  const int b[3] = {tr[0] * 4 + (l & 3), tr[1] * 4 + ((l >> 2) & 3), tr[2] * 4 + (l >> 4)};
  const int B = a.L.B, sh = a.L.sF, nF = a.L.nF;
  const size_t base = (size_t)pageToSlot(page, a.L.side) * (size_t)(nF >> 5);
  for (int cell0 = 0; cell0 < nF; cell0 += 64) {
    const int cell = cell0 + lane;
    const int x = b[0] * B + (cell & (B - 1)), y = b[1] * B + ((cell >> sh) & (B - 1)), z = b[2] * B + (cell >> (2 * sh));
    bool inside = false;
    for (int ci = 0; ci < a.colliderCount && !inside; ci++)
      inside = insideCollider(a.colliders[ci], a.colliderMasks, a.P.hf, x, y, z);
    const uint64_t bits = __ballot(inside);
    if (lane < 2) a.solids[base + (size_t)(cell0 >> 5) + lane] = (uint32_t)(bits >> (32 * lane));
  }
}
__device__ __forceinline__ void build_solids_body(const AllocArgs &a, const int vb) {
  if (!a.buildSolids) return;
  const int i = vb;
  if (a.solidsEvery) {
    if (i >= a.counts[0]) return;
    const int index = a.active[i];
    solidBrick(a, index, a.entries[index] & 0xffffff, (int)threadIdx.x);
  } else {
    if (i >= a.counts[1]) return;
    solidBrick(a, a.zeroList[2 * i], a.zeroList[2 * i + 1] & 0xffffff, (int)threadIdx.x);
  }
}
FP_KERNEL_PAIR(build_solids, 64, AllocArgs)


// A fresh slot is still, empty air (zeroVelocityBricks / zeroFieldBricks,
// fluid-zeroing.wgsl:76-105): its fields, expansion and velocity are cleared.
__device__ __forceinline__ void zero_fresh_body(const AllocArgs &a, const int vb, int chunks) {
  const Pools &p = a.p;
  const int i = vb / chunks, chunk = vb % chunks;
  if (i >= a.counts[1]) return;
  int slot = pageToSlot(a.zeroList[2 * i + 1] & 0xffffff, a.L.side);
  const int nF = a.L.nF, nV = a.L.nV;
  for (int c = chunk * 256 + threadIdx.x; c < nF; c += chunks * 256) {
    p.F[(size_t)slot * nF + c] = make_float4(0.f, 0.f, 0.f, 0.f);
    p.E[(size_t)slot * nF + c] = 0.f;
  }
  for (int c = chunk * 256 + threadIdx.x; c < 3 * nV; c += chunks * 256) p.V[(size_t)slot * nV * 3 + c] = 0.f;
  // zeroSmokeBricks (smoke.wgsl:187-223)
  if (a.L.sd > 1)
    for (int c = chunk * 256 + threadIdx.x; c < a.L.nS; c += chunks * 256) p.S[(size_t)slot * a.L.nS + c] = 0.f;
}
FP_KERNEL_PAIR1(zero_fresh, 256, AllocArgs, int, chunks)
// The same over the zeroed list only, walked by `grid` workgroups (a step zeroes few
// bricks), this one `block`; with `solids` the list's solids too, a wavefront a brick (the
// two are independent).
__device__ __forceinline__ void freshZero(const AllocArgs &a, int block, int grid, int chunks, int solids) {
  const int total = a.counts[1] * chunks;
  for (int v = block; v < total; v += grid) zero_fresh_body(a, v, chunks);
  if (!solids) return;
  const int waves = (int)blockDim.x >> 6, lane = (int)threadIdx.x & 63;
  for (int i = block * waves + ((int)threadIdx.x >> 6); i < a.counts[1]; i += grid * waves)
    solidBrick(a, a.zeroList[2 * i], a.zeroList[2 * i + 1] & 0xffffff, lane);
}
// Linking and the fresh bricks' zeroing in one launch: the first `linkBlocks` workgroups
// link (they write the brick records), the rest walk the zeroed list (they write the pools
// and solid bits), so neither reads what the other writes.
__global__ void __launch_bounds__(256) k_link_fresh(AllocArgs a, int linkBlocks, int chunks, int solids) {
  if ((int)blockIdx.x < linkBlocks) {
    link_bricks_body(a, (int)blockIdx.x);
    return;
  }
  freshZero(a, (int)blockIdx.x - linkBlocks, (int)gridDim.x - linkBlocks, chunks, solids);
}


}  // namespace fp
