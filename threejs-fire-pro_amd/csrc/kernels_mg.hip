// Sparse geometric multigrid for the pressure (PressureSolver.ts:276-402, pressure.wgsl).
// Brick levels halve the velocity brick down to 2^3 cells per pool slot; three tile levels
// then hold 4^3, 2^3 and one cell per tile slot. Each step computes one byte per cell per
// level: the cached open-side/solid topology plus whether the cell is a pressure unknown,
// so every sweep reads a single flag byte per cell.
#include "kernels.hpp"

namespace fp {

static __device__ __constant__ int OFF[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

struct MGCtx {
  Dir dir;
  Layout L;
  int side;
  bool hasBrick;
  Home home;
  bool hasTile;
  int tc[3];
  const int *tnb;  // the home tile's 27 neighbor slots (LDS, or the tile record itself)
};

__device__ __forceinline__ int levelShift(const Layout &L, int lv) {
  return lv < L.brickLevels ? L.sV - lv : L.sV + 2 - lv;
}

__device__ __forceinline__ int tileSlotAt(const MGCtx &m, int x, int y, int z) {
  if (m.hasTile) {
    int dx = x - m.tc[0], dy = y - m.tc[1], dz = z - m.tc[2];
    if ((unsigned)(dx + 1) <= 2u && (unsigned)(dy + 1) <= 2u && (unsigned)(dz + 1) <= 2u)
      return m.tnb[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
  }
  return findTile(m.dir, x, y, z);
}
__device__ __forceinline__ int brickEntry(const MGCtx &m, int bx, int by, int bz) {
  if (m.hasBrick) return entryAt(m.home, m.dir, bx, by, bz);
  int t = tileSlotAt(m, bx >> 2, by >> 2, bz >> 2);
  return t < 0 ? -1 : m.dir.entries[t * 64 + (((bz & 3) * 4 + (by & 3)) * 4 + (bx & 3))];
}

// cellAddress (pressure.wgsl:71-92).
__device__ __forceinline__ long addrAt(const MGCtx &m, const int p[3], int lv) {
  int sh = levelShift(m.L, lv), sd = 1 << sh;
  int ox = p[0] >> sh, oy = p[1] >> sh, oz = p[2] >> sh;
  int owner = lv < m.L.brickLevels ? entrySlot(brickEntry(m, ox, oy, oz), m.side) : tileSlotAt(m, ox, oy, oz);
  if (owner < 0) return -1;
  int lx = p[0] & (sd - 1), ly = p[1] & (sd - 1), lz = p[2] & (sd - 1);
  return (long)owner * sd * sd * sd + (lz * sd + ly) * sd + lx;
}

// addrAt of a brick-level cell within the home brick's 3^3 neighborhood, through the home
// slot table (the caller guarantees the neighborhood): no lookup and no branch on where the
// cell lies.
__device__ __forceinline__ long addrHome(const MGCtx &m, const int p[3], int lv) {
  const int sh = levelShift(m.L, lv), sd = 1 << sh, msk = sd - 1;
  const int rel = (((p[2] >> sh) - m.home.bz + 1) << 4) | (((p[1] >> sh) - m.home.by + 1) << 2) |
                  ((p[0] >> sh) - m.home.bx + 1);
  const int owner = m.home.slot[rel];
  const int local = (((p[2] & msk) << sh | (p[1] & msk)) << sh) | (p[0] & msk);
  return owner < 0 ? -1 : (long)owner * sd * sd * sd + local;
}
// addrAt of a tile-level cell within the context tile's 3^3 neighborhood, through its 27
// neighbor slots (the caller guarantees the neighborhood).
__device__ __forceinline__ long addrTileNear(const MGCtx &m, const int p[3], int lv) {
  const int sh = levelShift(m.L, lv), sd = 1 << sh, msk = sd - 1;
  const int owner = m.tnb[((p[2] >> sh) - m.tc[2] + 1) * 9 + ((p[1] >> sh) - m.tc[1] + 1) * 3 + ((p[0] >> sh) - m.tc[0] + 1)];
  const int local = (((p[2] & msk) << sh | (p[1] & msk)) << sh) | (p[0] & msk);
  return owner < 0 ? -1 : (long)owner * sd * sd * sd + local;
}
// addrAt, or for NEAR a cell within one cell of the context's cube: through the home brick's
// slots at a brick level, the context tile's neighbors at a tile level.
template <bool NEAR>
__device__ __forceinline__ long addrOf(const MGCtx &m, const int p[3], int lv) {
  if (!NEAR) return addrAt(m, p, lv);
  return lv < m.L.brickLevels ? addrHome(m, p, lv) : addrTileNear(m, p, lv);
}

// activeCell (pressure.wgsl:101-132).
__device__ __forceinline__ bool activeCell(const MGCtx &m, const Params &P, const int p[3], int lv) {
  const Layout &L = m.L;
  if (lv >= L.brickLevels) {
    int n = 1 << (lv - L.brickLevels);
    for (int z = 0; z < n; z++)
      for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
          int bx = p[0] * n + x, by = p[1] * n + y, bz = p[2] * n + z;
          int t = tileSlotAt(m, bx >> 2, by >> 2, bz >> 2);
          if (t < 0) return false;
          int id = t * 64 + (((bz & 3) * 4 + (by & 3)) * 4 + (bx & 3));
          if (m.dir.entries[id] < 0 || m.dir.records[(size_t)id * RECORD_WORDS + PRESSURE_FLAG] == 0) return false;
        }
    return true;
  }
  int span = 1 << lv;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = (p[a] * span - 1) >> L.sV;
    hi[a] = (p[a] * span + span - 1) >> L.sV;
  }
  if (P.ground) lo[1] = lo[1] > 0 ? lo[1] : 0;
  // At most two bricks a side (a brick-level cell spans at most half a brick), beside the
  // home brick of a brick context's own cells; the floor only raises lo, and a box it leaves
  // non-empty still starts at most one brick below home.
  if (m.hasBrick) return bricksPresent<true>(m.home, m.dir, lo, hi);
  for (int z = lo[2]; z <= hi[2]; z++)
    for (int y = lo[1]; y <= hi[1]; y++)
      for (int x = lo[0]; x <= hi[0]; x++)
        if (brickEntry(m, x, y, z) < 0) return false;
  return true;
}

// solidAt for the pressure solver: through the directory, not a fluid home brick.
__device__ __forceinline__ bool solidAtMG(const MGCtx &m, const Params &P, const uint32_t *solids, float wx, float wy, float wz) {
  int x = (int)floorf(wx * P.invHf + 0.25f), y = (int)floorf(wy * P.invHf + 0.25f), z = (int)floorf(wz * P.invHf + 0.25f);
  if (P.ground && y < 0) return true;
  if (P.colliders == 0 || !inSolidBox(P, x, y, z)) return false;
  const int sF = m.L.sF;
  int s = entrySlot(brickEntry(m, x >> sF, y >> sF, z >> sF), m.side);
  if (s < 0) return false;
  int mk = m.L.B - 1;
  uint32_t bit = (uint32_t)(((((z & mk) << sF) | (y & mk)) << sF) | (x & mk));
  return ((solids[(size_t)s * (m.L.nF >> 5) + (bit >> 5)] >> (bit & 31u)) & 1u) != 0u;
}
__device__ __forceinline__ bool blockedAt(const MGCtx &m, const Params &P, const uint32_t *solids, const int p[3], int lv) {
  float scale = (float)(1 << lv);
  return solidAtMG(m, P, solids, ((float)p[0] + 0.5f) * P.hv * scale, ((float)p[1] + 0.5f) * P.hv * scale,
                   ((float)p[2] + 0.5f) * P.hv * scale);
}
// cellTopology (pressure.wgsl:215-226).
__device__ __forceinline__ uint8_t cellTopology(const MGCtx &m, const Params &P, const uint32_t *solids, const int p[3], int lv) {
  if (blockedAt(m, P, solids, p, lv)) return STORED_SOLID;
  uint8_t flags = 0;
  for (int sd = 0; sd < 6; sd++) {
    int q[3] = {p[0] + OFF[sd][0], p[1] + OFF[sd][1], p[2] + OFF[sd][2]};
    if (!blockedAt(m, P, solids, q, lv)) flags |= (uint8_t)(1u << sd);
  }
  return flags;
}

// Level sides are powers of two: the cell's offsets are masks and shifts.
__device__ __forceinline__ void levelCell(int home0, int home1, int home2, int sd, int l, int p[3]) {
  const int k = __builtin_ctz((unsigned)sd), msk = sd - 1;
  p[0] = home0 * sd + (l & msk);
  p[1] = home1 * sd + ((l >> k) & msk);
  p[2] = home2 * sd + (l >> (2 * k));
}

// Home set-up for a block over one active brick.
#define MG_BRICK_HOME(ID)                          \
  __shared__ int rec[HOME_WORDS];                          \
  MGCtx c;                                         \
  c.dir = m.dir;                                   \
  c.L = m.L;                                       \
  c.side = m.L.side;                               \
  c.hasBrick = true;                               \
  brickTile(c, ID);                                \
  loadHome(m.dir, ID, rec, c.home, m.L.side);

// Home set-up for a block over one tile slot; returns if the slot is free.
#define MG_TILE_HOME(T)                                                         \
  __shared__ int tnb[27];                                                       \
  const int *trec = m.dir.tileRec + (size_t)(T) * TILE_RECORD_WORDS;            \
  if (trec[3] == 0) return;                                                     \
  if (threadIdx.x < 27) tnb[threadIdx.x] = trec[4 + threadIdx.x];               \
  __syncthreads();                                                              \
  MGCtx c;                                                                      \
  c.dir = m.dir;                                                                \
  c.L = m.L;                                                                    \
  c.side = m.L.side;                                                            \
  c.hasBrick = false;                                                           \
  c.hasTile = true;                                                             \
  c.tc[0] = trec[0];                                                            \
  c.tc[1] = trec[1];                                                            \
  c.tc[2] = trec[2];                                                            \
  c.tnb = tnb;

// Brick-level and tile-level sweeps touch few cells per brick at the coarse levels (64,
// 8 or 1 per velocity brick of 4^3), so a workgroup takes a group of consecutive items
// (active bricks or tile slots) with `cells` lanes each: group g covers items g*k .. g*k+k-1
// with k = lanes / cells (at most MG_GROUP); lane t works on cell t % cells of item
// t / cells. With cells >= lanes, k = 1 and the lanes stride over the item's cells.
__device__ __forceinline__ int groupSize(int cells, int lanes) {
  return cells >= lanes ? 1 : (lanes / cells < MG_GROUP ? lanes / cells : MG_GROUP);
}
struct Group {
  int k, g, j, first, step;  // items per group, group, this lane's item in it, first cell, cell stride
};
__device__ __forceinline__ Group groupOf(int g, int cells) {
  Group r;
  r.k = groupSize(cells, (int)blockDim.x);
  r.g = g;
  if (r.k == 1) {
    r.j = 0;
    r.first = (int)threadIdx.x;
    r.step = (int)blockDim.x;
  } else {
    r.j = (int)threadIdx.x / cells;
    r.first = (int)threadIdx.x % cells;
    r.step = cells;
    if (r.j >= r.k) r.j = -1;  // lanes past the group's k * cells
  }
  return r;
}
// The group's bricks' 27 neighbor entries and their pool slots, in LDS (all lanes take
// part); this lane's brick's home.
__device__ __forceinline__ void loadGroupHomes(const Dir &dir, const int *active, int count, const Group &gr,
                                               int (*recs)[HOME_WORDS], int side) {
  for (int e = threadIdx.x; e < gr.k * 27; e += blockDim.x) {
    const int jj = e / 27, w = e - jj * 27;
    const int b = gr.g * gr.k + jj;
    if (b < count) {
      const int r = dir.records[(size_t)active[b] * RECORD_WORDS + w];
      recs[jj][w] = r;
      recs[jj][27 + homeSlotWord(w)] = entrySlot(r, side);
    }
  }
  __syncthreads();
}
__device__ __forceinline__ Home homeOf(const Dir &dir, int id, const int *rec) {
  Home h;
  const int *tr = dir.tileRec + (size_t)(id >> 6) * TILE_RECORD_WORDS;
  const int l = id & 63;
  h.bx = tr[0] * 4 + (l & 3);
  h.by = tr[1] * 4 + ((l >> 2) & 3);
  h.bz = tr[2] * 4 + (l >> 4);
  h.rec = rec;
  h.slot = rec + 27;
  return h;
}
// A brick context's tile neighborhood: the tile holding brick `id` and its 26 neighbor
// slots, read in place from the tile record, so that cells of the tile levels near the brick
// resolve their tile slot without a hash probe.
__device__ __forceinline__ void brickTile(MGCtx &c, int id) {
  const int *tr = c.dir.tileRec + (size_t)(id >> 6) * TILE_RECORD_WORDS;
  c.hasTile = true;
  c.tc[0] = tr[0];
  c.tc[1] = tr[1];
  c.tc[2] = tr[2];
  c.tnb = tr + 4;
}
// A group of active bricks with `cells` cells each at the level swept: declares `gr`, the
// brick index `bi` of this lane (valid when `mine`), and with HOME its MGCtx `c`.
#define MG_BRICK_GROUP(CELLS)                                                   \
  const int count_ = m.counts[0];                                              \
  const int k_ = groupSize((CELLS), (int)blockDim.x);                          \
  const int g_ = xcdItem(vb, (count_ + k_ - 1) / k_);                          \
  if (g_ < 0) return;                                                          \
  const Group gr = groupOf(g_, (CELLS));                                       \
  const int bi = gr.g * gr.k + gr.j;                                           \
  const bool mine = gr.j >= 0 && bi < count_;
#define MG_GROUP_HOME                                                           \
  __shared__ int recs_[MG_GROUP][HOME_WORDS];                                          \
  loadGroupHomes(m.dir, m.active, count_, gr, recs_, m.L.side);                \
  if (!mine) return;                                                           \
  const int id = m.active[bi];                                                 \
  MGCtx c;                                                                     \
  c.dir = m.dir;                                                               \
  c.L = m.L;                                                                   \
  c.side = m.L.side;                                                           \
  c.hasBrick = true;                                                           \
  brickTile(c, id);                                                            \
  c.home = homeOf(m.dir, id, recs_[gr.j]);                                     \
  const int *rec = recs_[gr.j];
// A group of tile slots with `cells` cells each: declares `gr`, this lane's tile `t` and,
// for a live slot, its MGCtx `c` (27 neighbor slots in LDS).
#define MG_TILE_GROUP(CELLS)                                                    \
  const int k_ = groupSize((CELLS), (int)blockDim.x);                          \
  if (vb * k_ >= m.dir.capacity) return;                                       \
  const Group gr = groupOf(vb, (CELLS));                                       \
  const int t = gr.g * gr.k + gr.j;                                            \
  __shared__ int tnbs_[MG_GROUP][27];                                          \
  for (int e = threadIdx.x; e < gr.k * 27; e += blockDim.x) {                  \
    const int jj = e / 27, w = e - jj * 27, tt = gr.g * gr.k + jj;             \
    if (tt < m.dir.capacity) tnbs_[jj][w] = m.dir.tileRec[(size_t)tt * TILE_RECORD_WORDS + 4 + w]; \
  }                                                                            \
  __syncthreads();                                                             \
  if (gr.j < 0 || t >= m.dir.capacity) return;                                 \
  const int *trec = m.dir.tileRec + (size_t)t * TILE_RECORD_WORDS;             \
  if (trec[3] == 0) return;                                                    \
  MGCtx c;                                                                     \
  c.dir = m.dir;                                                               \
  c.L = m.L;                                                                   \
  c.side = m.L.side;                                                           \
  c.hasBrick = false;                                                          \
  c.hasTile = true;                                                            \
  c.tc[0] = trec[0];                                                           \
  c.tc[1] = trec[1];                                                           \
  c.tc[2] = trec[2];                                                           \
  c.tnb = tnbs_[gr.j];

// buildBrickTopology (pressure.wgsl:259-278) for every active brick (when the geometry
// changed), or for the bricks of the zeroed list that hold a slot.
__device__ __forceinline__ void topology_brick(const MGArgs &m, const int id) {
  if (m.dir.entries[id] < 0) return;
  MG_BRICK_HOME(id)
  int slot = entrySlot(rec[13], c.side);
  // Every brick level's cells as one index space: a lane's cells of different levels are
  // independent, so one pass of the workgroup covers them all.
  int total = 0;
  for (int lv = 0; lv < m.L.brickLevels; lv++) total += m.lv[lv].side * m.lv[lv].side * m.lv[lv].side;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int lv = 0, l = i, n = m.lv[0].side * m.lv[0].side * m.lv[0].side;
    while (l >= n) {
      l -= n;
      lv++;
      n = m.lv[lv].side * m.lv[lv].side * m.lv[lv].side;
    }
    int p[3];
    levelCell(c.home.bx, c.home.by, c.home.bz, m.lv[lv].side, l, p);
    m.lv[lv].topo[(size_t)slot * n + l] = cellTopology(c, m.P, m.solids, p, lv);
  }
}
// Whether a collider can make any cell that a topology of brick `id` samples solid: the
// brick's field cells and one brick around them against the colliders' box of field cells.
// Elsewhere a brick's topology depends on the floor alone.
__device__ __forceinline__ bool nearSolids(const MGArgs &m, int id) {
  if (m.P.colliders == 0) return false;
  const int *tr = m.dir.tileRec + (size_t)(id >> 6) * TILE_RECORD_WORDS;
  const int l = id & 63, B = m.L.B;
  const int b[3] = {tr[0] * 4 + (l & 3), tr[1] * 4 + ((l >> 2) & 3), tr[2] * 4 + (l >> 4)};
  for (int a = 0; a < 3; a++)
    if (b[a] * B + 2 * B - 1 < m.P.solidLo[a] || b[a] * B - B > m.P.solidHi[a]) return false;
  return true;
}
__device__ __forceinline__ void topology_bricks_body(const MGArgs &m, const int vb, int fromZeroList) {
  // With the geometry changed, the pass over every active brick (launched first) has
  // covered the zeroed list's bricks that hold a slot.
  if (fromZeroList ? m.topologyAll : !m.topologyAll) return;
  const int bi = xcdItem(vb, fromZeroList ? m.counts[1] + m.counts[5] : m.counts[0]);
  if (bi < 0) return;
  topology_brick(m, fromZeroList ? m.zeroList[2 * bi] : m.active[bi]);
}
FP_KERNEL_PAIR1(topology_bricks, 256, MGArgs, int, fromZeroList)
// The same over the zeroed list only, walked by a fixed grid: a step allocates few bricks,
// and a grid sized for the pool would be almost all empty workgroups.
__global__ void __launch_bounds__(256) k_fresh_topology(MGArgs m) {
  const int count = m.counts[1] + m.counts[5];
  for (int bi = (int)blockIdx.x; bi < count; bi += (int)gridDim.x) {
    topology_brick(m, m.zeroList[2 * bi]);
    __syncthreads();  // the LDS record is loaded again for the next brick
  }
}

// When only the tile set changed: the bricks of the listed tile slots (64 workgroups a tile)
// that hold a slot and lie near the colliders (nearSolids).
__global__ void __launch_bounds__(256) k_near_topology(MGArgs m, const int *tiles, int count) {
  const int vb = (int)blockIdx.x;
  if (vb >= count * 64) return;
  const int id = tiles[vb >> 6] * 64 + (vb & 63);
  if (m.dir.entries[id] < 0 || !nearSolids(m, id)) return;
  topology_brick(m, id);
}

// buildTileTopology (pressure.wgsl:281-305) of one tile level (`only`), or of every tile
// level (`only` < 0), for every tile slot.
__device__ __forceinline__ void topology_tiles_body(const MGArgs &m, const int vb, int only) {
  if (!m.topologyAll && !m.topologyTiles) return;
  const int t = vb;
  if (t >= m.dir.capacity) return;
  const int first = only < 0 ? m.L.brickLevels : only, last = only < 0 ? m.bottom : only;
  const int *trec0 = m.dir.tileRec + (size_t)t * TILE_RECORD_WORDS;
  if (trec0[3] == 0) {
    for (int lv = first; lv <= last; lv++) {
      const int sd = m.lv[lv].side, n = sd * sd * sd;
      for (int l = threadIdx.x; l < n; l += blockDim.x) m.lv[lv].topo[(size_t)t * n + l] = 0;
    }
    return;
  }
  MG_TILE_HOME(t)
  // The levels' cells as one index space, one pass of the workgroup (as topology_brick).
  int total = 0;
  for (int lv = first; lv <= last; lv++) total += m.lv[lv].side * m.lv[lv].side * m.lv[lv].side;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int lv = first, l = i, n = m.lv[first].side * m.lv[first].side * m.lv[first].side;
    while (l >= n) {
      l -= n;
      lv++;
      n = m.lv[lv].side * m.lv[lv].side * m.lv[lv].side;
    }
    int p[3];
    levelCell(c.tc[0], c.tc[1], c.tc[2], m.lv[lv].side, l, p);
    m.lv[lv].topo[(size_t)t * n + l] = cellTopology(c, m.P, m.solids, p, lv);
  }
}
FP_KERNEL_PAIR1(topology_tiles, 128, MGArgs, int, lv)

// buildTileTopology for the listed tile slots only (a workgroup each), when only the tile set
// changed: the created tiles and the tiles near the colliders.
__global__ void __launch_bounds__(128) k_topology_tiles_list(MGArgs m, const int *tiles, int count) {
  if ((int)blockIdx.x >= count) return;
  topology_tiles_body(m, tiles[blockIdx.x], -1);
}

// The three passes of a step whose tile set changed, in one launch of 128-lane workgroups:
// the zeroed list's bricks (the first `fresh` workgroups, walking it as k_fresh_topology),
// the bricks of the first `nearCount` listed tiles (64 workgroups a tile, as
// k_near_topology) and the tile levels of the next `tileCount` (as k_topology_tiles_list).
// Every part computes topology from the solids and the tables alone, so they are
// independent; a brick both fresh and near gets the same bytes from both.
__global__ void __launch_bounds__(128) k_moved_topology(MGArgs m, int fresh, const int *tiles, int nearCount,
                                                        int tileCount) {
  int vb = (int)blockIdx.x;
  if (vb < fresh) {
    const int count = m.counts[1] + m.counts[5];
    for (int bi = vb; bi < count; bi += fresh) {
      topology_brick(m, m.zeroList[2 * bi]);
      __syncthreads();  // the LDS record is loaded again for the next brick
    }
    return;
  }
  vb -= fresh;
  if (vb < nearCount * 64) {
    const int id = tiles[vb >> 6] * 64 + (vb & 63);
    if (m.dir.entries[id] < 0 || !nearSolids(m, id)) return;
    topology_brick(m, id);
    return;
  }
  vb -= nearCount * 64;
  if (vb < tileCount) topology_tiles_body(m, tiles[nearCount + vb], -1);
}

// The step's flags of the brick levels: topology plus the unknown bit.
__device__ __forceinline__ void eflags_bricks_body(const MGArgs &m, const int vb) {
  const int bi = xcdItem(vb, m.counts[0]);
  if (bi < 0) return;
  const int id = m.active[bi];
  MG_BRICK_HOME(id)
  int slot = entrySlot(rec[13], c.side);
  for (int lv = 0; lv < m.L.brickLevels; lv++) {
    int sd = m.lv[lv].side, n = sd * sd * sd;
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
      int p[3];
      levelCell(c.home.bx, c.home.by, c.home.bz, sd, l, p);
      size_t a = (size_t)slot * n + l;
      m.lv[lv].eflag[a] = (uint8_t)(m.lv[lv].topo[a] | (activeCell(c, m.P, p, lv) ? UNKNOWN : 0));
    }
  }
}
FP_KERNEL_PAIR(eflags_bricks, 256, MGArgs)
// The step's flags of every tile level of one tile slot, from one wavefront (64 lanes, a
// lane per brick of the tile, (z * 4 + y) * 4 + x). A cell of the first tile level is one
// brick, unknown when the brick has a slot and takes part in the solve; a cell of a
// coarser tile level covers its eight children's bricks, so it is unknown when all of
// them are (the brick test of activeCell over the cell's bricks).
__device__ __forceinline__ void eflags_tiles_body(const MGArgs &m, const int vb) {
  const int t = vb;
  if (t >= m.dir.capacity) return;
  if (m.dir.tileRec[(size_t)t * TILE_RECORD_WORDS + 3] == 0) return;
  const int l = (int)threadIdx.x, id = t * 64 + l;
  const bool part = m.dir.entries[id] >= 0 && m.dir.records[(size_t)id * RECORD_WORDS + PRESSURE_FLAG] != 0;
  const uint64_t bricks = __ballot(part);
  for (int lv = m.L.brickLevels; lv <= m.bottom; lv++) {
    const int k = 2 - (lv - m.L.brickLevels);  // log2 of the level's side in the tile
    if (k < 0) break;
    const int sd = 1 << k, n = 1 << (3 * k), span = 4 >> k;
    if (l >= n) continue;
    const int c[3] = {(l & (sd - 1)) * span, ((l >> k) & (sd - 1)) * span, (l >> (2 * k)) * span};
    bool all = true;
    for (int z = 0; z < span; z++)
      for (int y = 0; y < span; y++)
        for (int x = 0; x < span; x++)
          all = all && ((bricks >> (((c[2] + z) * 4 + c[1] + y) * 4 + c[0] + x)) & 1u);
    const size_t a = (size_t)t * n + l;
    m.lv[lv].eflag[a] = (uint8_t)(m.lv[lv].topo[a] | (all ? UNKNOWN : 0));
  }
}
FP_KERNEL_PAIR(eflags_tiles, 64, MGArgs)

// A level's arrays as one launch sees them: its current iterate p, the other array q.
template <class T> struct LevelView {
  T *p, *q, *rhs;
  const uint8_t *topo, *eflag;
  int side;
};
template <class T> __device__ __forceinline__ LevelView<T> levelView(const MGArgsT<T> &m, int L, int flips) {
  const LevelT<T> &l = m.lv[L];
  const bool f = ((flips >> L) & 1) != 0;
  return LevelView<T>{f ? l.q : l.p, f ? l.p : l.q, l.rhs, l.topo, l.eflag, l.side};
}

template <class St, bool HOME = false>
__device__ __forceinline__ float readAt(const MGCtx &c, const typename St::T *v, const int p[3], int lv) {
  long a = addrOf<HOME>(c, p, lv);
  return a < 0 ? 0.f : St::ld(v[a]);
}
__device__ __forceinline__ float openCount(int flags, int sd) {
  return (float)((flags >> sd) & 1) + (float)((flags >> (sd + 1)) & 1);
}
__device__ __forceinline__ float diagonal(float w, int flags) {
  return w * openCount(flags, 0) + w * openCount(flags, 2) + w * openCount(flags, 4);
}
// stencil (pressure.wgsl:376-402): the neighbor sum over open sides. `a` is the cell's own
// index in the level array, whose slots hold cubes `sd` cells a side: a neighbor inside the
// same cube is at a fixed offset from `a`; only neighbors across a face resolve their brick
// or tile. The six neighbors are read whatever the flags say (their addresses do not depend
// on them), so the reads overlap the cell's own flag read; closed sides then count zero.
template <class St>
__device__ __forceinline__ void neighbors(const MGCtx &c, const typename St::T *pv, const int p[3], int lv, long a,
                                          int sd, float n[6]) {
  const int m = sd - 1;
  const int local[3] = {p[0] & m, p[1] & m, p[2] & m};
  for (int k = 0; k < 6; k++) {
    const int lx = local[0] + OFF[k][0], ly = local[1] + OFF[k][1], lz = local[2] + OFF[k][2];
    if ((unsigned)lx < (unsigned)sd && (unsigned)ly < (unsigned)sd && (unsigned)lz < (unsigned)sd) {
      n[k] = St::ld(pv[a + OFF[k][0] + (OFF[k][1] + OFF[k][2] * sd) * sd]);
    } else {
      int q[3] = {p[0] + OFF[k][0], p[1] + OFF[k][1], p[2] + OFF[k][2]};
      n[k] = readAt<St>(c, pv, q, lv);
    }
  }
}
__device__ __forceinline__ float openSum(const float n[6], int flags, float w) {
  float o[6];
  for (int k = 0; k < 6; k++) o[k] = ((flags >> k) & 1) ? n[k] : 0.f;
  return w * (o[0] + o[1]) + w * (o[2] + o[3]) + w * (o[4] + o[5]);
}
template <class St>
__device__ __forceinline__ float stencilSum(const MGCtx &c, const typename St::T *pv, const int p[3], int flags, int lv,
                                            float w, long a, int sd) {
  float n[6];
  neighbors<St>(c, pv, p, lv, a, sd, n);
  return openSum(n, flags, w);
}

// The step's flags of every level and level 0's first iterate (initializePressure,
// pressure.wgsl:485-505) in one launch, for the single V-cycle. Workgroups below
// `brickGroups` take groups of active bricks with one lane per level-0 cell (the grouping of
// the level-0 sweeps): each lane sets the flags of its cells of every brick level, as
// eflags_bricks does, and the first iterate of its level-0 cells from the flags it has just
// set. The workgroups after them take four tile slots each, a wavefront per slot, as
// eflags_tiles does.
__global__ void __launch_bounds__(256) k_mg_flags_init(MGArgs m, int brickGroups) {
  const int vb = (int)blockIdx.x;
  if (vb >= brickGroups) {
    const int t = (vb - brickGroups) * 4 + (int)(threadIdx.x >> 6);
    if (t >= m.dir.capacity) return;
    if (m.dir.tileRec[(size_t)t * TILE_RECORD_WORDS + 3] == 0) return;
    const int l = (int)(threadIdx.x & 63), id = t * 64 + l;
    const bool part = m.dir.entries[id] >= 0 && m.dir.records[(size_t)id * RECORD_WORDS + PRESSURE_FLAG] != 0;
    const uint64_t bricks = __ballot(part);
    for (int lv = m.L.brickLevels; lv <= m.bottom; lv++) {
      const int k = 2 - (lv - m.L.brickLevels);
      if (k < 0) break;
      const int sd = 1 << k, n = 1 << (3 * k), span = 4 >> k;
      if (l >= n) continue;
      const int cc[3] = {(l & (sd - 1)) * span, ((l >> k) & (sd - 1)) * span, (l >> (2 * k)) * span};
      bool all = true;
      for (int z = 0; z < span; z++)
        for (int y = 0; y < span; y++)
          for (int x = 0; x < span; x++)
            all = all && ((bricks >> (((cc[2] + z) * 4 + cc[1] + y) * 4 + cc[0] + x)) & 1u);
      const size_t a = (size_t)t * n + l;
      m.lv[lv].eflag[a] = (uint8_t)(m.lv[lv].topo[a] | (all ? UNKNOWN : 0));
    }
    return;
  }
  const int n0 = m.lv[0].side * m.lv[0].side * m.lv[0].side;
  MG_BRICK_GROUP(n0)
  MG_GROUP_HOME
  const int slot = entrySlot(rec[13], c.side);
  for (int lv = 0; lv < m.L.brickLevels; lv++) {
    const int sd = m.lv[lv].side, n = sd * sd * sd;
    for (int l = gr.first; l < n; l += gr.step) {
      int p[3];
      levelCell(c.home.bx, c.home.by, c.home.bz, sd, l, p);
      const size_t a = (size_t)slot * n + l;
      const int e = m.lv[lv].topo[a] | (activeCell(c, m.P, p, lv) ? UNKNOWN : 0);
      m.lv[lv].eflag[a] = (uint8_t)e;
      if (lv == 0) {
        float first = 0.f;
        if (!(e & STORED_SOLID) && (e & UNKNOWN))
          first = mixv(0.f, (0.f - m.lv[0].rhs[a]) / fmaxv(diagonal(m.weight, e & 63), 1e-12f), 0.8f);
        m.lv[0].p[a] = first;
      }
    }
  }
}

// The six neighbors of brick-level cell p (level cubes sd = 2^sh cells a side) through the
// home brick's slot table: a neighbor lies in the home brick or one beside it, so every lane
// reads a slot from LDS and loads (from slot 0 where there is none, then zero), with no
// branch on where the neighbor lies.
template <class St>
__device__ __forceinline__ void brickNeighbors(const MGCtx &c, const typename St::T *pv, const int p[3], int sh,
                                               float n[6]) {
  constexpr int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};  // OFF
  const int sd = 1 << sh, m = sd - 1, cells = sd * sd * sd;
  const int home[3] = {c.home.bx, c.home.by, c.home.bz};
#pragma unroll
  for (int k = 0; k < 6; k++) {
    int rel = 0, local = 0;
#pragma unroll
    for (int ax = 2; ax >= 0; ax--) {
      const int q = p[ax] + off[k][ax];
      rel |= ((q >> sh) - home[ax] + 1) << (2 * ax);
      local = local * sd + (q & m);
    }
    const int slot = c.home.slot[rel];
    const float v = St::ld(pv[(size_t)(slot < 0 ? 0 : slot) * cells + local]);
    n[k] = slot < 0 ? 0.f : v;
  }
}
// relaxCell (pressure.wgsl:468-479): weighted Jacobi, omega = 0.8, p to q. BRICK: a brick
// level (neighbors through brickNeighbors).
template <class St, bool BRICK = false, bool TILE = false>
__device__ __forceinline__ void relaxCell(const MGCtx &c, const LevelView<typename St::T> &lv, int level, const int p[3],
                                          size_t a, float w) {
  // Every read is issued before the flags are looked at.
  const int e = lv.eflag[a];
  const float own = St::ld(lv.p[a]), rhs = St::ld(lv.rhs[a]);
  float n[6];
  if (BRICK) {
    brickNeighbors<St>(c, lv.p, p, levelShift(c.L, level), n);
  } else if (TILE) {
    // A tile level's neighbors lie in the context tile or one beside it.
    constexpr int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};  // OFF
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const int q[3] = {p[0] + off[k][0], p[1] + off[k][1], p[2] + off[k][2]};
      const long qa = addrTileNear(c, q, level);
      const float v = St::ld(lv.p[qa < 0 ? 0 : qa]);
      n[k] = qa < 0 ? 0.f : v;
    }
  } else {
    neighbors<St>(c, lv.p, p, level, (long)a, lv.side, n);
  }
  if ((e & STORED_SOLID) || !(e & UNKNOWN)) {
    lv.q[a] = St::st(0.f);
    return;
  }
  const int flags = e & 63;
  const float sum = openSum(n, flags, w);
  const float cand = (sum - rhs) / fmaxv(diagonal(w, flags), 1e-12f);
  lv.q[a] = St::st(mixv(own, cand, 0.8f));
}

// The single and batched kernels of a solve stage; a simulation whose defect correction
// has finished is skipped.
#define FP_MG_PAIR(NAME, BOUNDS)                                                                          \
  template <class St> __global__ void __launch_bounds__(BOUNDS) k_##NAME(MGArgsT<typename St::T> a, int code) { \
    if (dcFinished(a)) return;                                                                            \
    NAME##_body<St>(a, (int)blockIdx.x, code);                                                            \
  }                                                                                                       \
  template <class St>                                                                                     \
  __global__ void __launch_bounds__(BOUNDS)                                                               \
      k_##NAME##_batch(const MGArgsT<typename St::T> *args, const int *starts, int n, int code) {          \
    FP_BATCH_LOOP(NAME, if (!dcFinished(args[sim])) NAME##_body<St>(args[sim], v - starts[sim], code))    \
  }

// initializePressure (pressure.wgsl:485-505): level 0's first iterate from zero.
template <class St> __device__ __forceinline__ void mg_init_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const auto lv = levelView(m, 0, code >> 8);
  const int sd = lv.side, n = sd * sd * sd;
  MG_BRICK_GROUP(n)
  if (!mine) return;
  const int id = m.active[bi];
  const int slot = entrySlot(m.dir.entries[id], m.L.side);
  for (int l = gr.first; l < n; l += gr.step) {
    size_t a = (size_t)slot * n + l;
    int e = lv.eflag[a];
    float first = 0.f;
    if (!(e & STORED_SOLID) && (e & UNKNOWN))
      first = mixv(0.f, (0.f - St::ld(lv.rhs[a])) / fmaxv(diagonal(m.weight, e & 63), 1e-12f), 0.8f);
    lv.p[a] = St::st(first);
  }
}
FP_MG_PAIR(mg_init, 256)

template <class St>
__device__ __forceinline__ void mg_relax_bricks_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int level = code & 255;
  const auto lv = levelView(m, level, code >> 8);
  const int sd = lv.side, n = sd * sd * sd;
  MG_BRICK_GROUP(n)
  MG_GROUP_HOME
  const int slot = entrySlot(rec[13], c.side);
  for (int l = gr.first; l < n; l += gr.step) {
    int p[3];
    levelCell(c.home.bx, c.home.by, c.home.bz, sd, l, p);
    relaxCell<St, true>(c, lv, level, p, (size_t)slot * n + l, m.weight);
  }
}
FP_MG_PAIR(mg_relax_bricks, 256)
template <class St>
__device__ __forceinline__ void mg_relax_tiles_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int level = code & 255;
  const auto lv = levelView(m, level, code >> 8);
  const int sd = lv.side, n = sd * sd * sd;
  MG_TILE_GROUP(n)
  for (int l = gr.first; l < n; l += gr.step) {
    int p[3];
    levelCell(c.tc[0], c.tc[1], c.tc[2], sd, l, p);
    relaxCell<St, false, true>(c, lv, level, p, (size_t)t * n + l, m.weight);
  }
}
FP_MG_PAIR(mg_relax_tiles, 64)

// restrictCell (pressure.wgsl:601-647): the coarse right-hand side is half the sum of the
// eight children's residuals; the coarse first iterate is one Jacobi step from zero.
// Every fine level's cube (a brick's or a tile's cells of that level) has an even side, so
// the eight children share one cube: one owner lookup covers them, and the 24 cells beside
// their 2^3 block resolve their cube once per face. All reads are issued before any flag is
// tested; the residuals are then added in child order, as one child at a time would.
// restrictResidual returns the coarse right-hand side (before storage rounding) and whether
// every child is an unknown; restrictCell stores it and the first iterate.
template <class St, bool HOME = false>
__device__ __forceinline__ float restrictResidual(const MGCtx &c, const MGArgsT<typename St::T> &m, int level,
                                                  int flips, const int id[3], bool &every) {
  const auto fine = levelView(m, level - 1, flips);
  const int sd = fine.side, msk = sd - 1;
  const int p0[3] = {id[0] * 2, id[1] * 2, id[2] * 2};
  const long a0 = addrOf<HOME>(c, p0, level - 1);
  const int stride[3] = {1, sd, sd * sd};
  // Face f (+x, -x, +y, -y, +z, -z) holds four cells, j = u + 2 v over its two other axes
  // in axis order.
  float side[6][4];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const int ax = f >> 1, step = (f & 1) ? -1 : 2;
    const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;
    long qa;
    if (HOME) {
      int q[3] = {p0[0], p0[1], p0[2]};
      q[ax] += step;
      qa = addrOf<true>(c, q, level - 1);
    } else if ((unsigned)((p0[ax] & msk) + step) < (unsigned)sd) {
      qa = a0 < 0 ? -1 : a0 + step * stride[ax];
    } else {
      int q[3] = {p0[0], p0[1], p0[2]};
      q[ax] += step;
      qa = addrAt(c, q, level - 1);
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
      side[f][j] = qa < 0 ? 0.f : St::ld(fine.p[qa + (j & 1) * stride[u] + (j >> 1) * stride[v]]);
  }
  float own[8], rhs8[8];
  int e[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const long a = a0 + (k & 1) + ((k >> 1) & 1) * stride[1] + (k >> 2) * stride[2];
    e[k] = a0 < 0 ? 0 : fine.eflag[a];
    own[k] = a0 < 0 ? 0.f : St::ld(fine.p[a]);
    rhs8[k] = a0 < 0 ? 0.f : St::ld(fine.rhs[a]);
  }
  float residual = 0.f;
  every = true;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int x = k & 1, y = (k >> 1) & 1, z = k >> 2;
    const bool unknown = (e[k] & UNKNOWN) != 0;
    every = every && unknown;
    if (!unknown || (e[k] & STORED_SOLID)) continue;
    const int flags = e[k] & 63;
    const float n[6] = {x ? side[0][y + 2 * z] : own[k + 1], x ? own[k - 1] : side[1][y + 2 * z],
                        y ? side[2][x + 2 * z] : own[k + 2], y ? own[k - 2] : side[3][x + 2 * z],
                        z ? side[4][x + 2 * y] : own[k + 4], z ? own[k - 4] : side[5][x + 2 * y]};
    const float sum = openSum(n, flags, m.weight);
    const float dg = diagonal(m.weight, flags);
    residual += rhs8[k] - (sum - dg * own[k]);
  }
  return residual * 0.5f;
}
template <class St, bool HOME = false>
__device__ __forceinline__ void restrictCell(const MGCtx &c, const MGArgsT<typename St::T> &m, int level, int flips,
                                             const int id[3]) {
  const auto coarse = levelView(m, level, flips);
  bool every;
  const float rhs = restrictResidual<St, HOME>(c, m, level, flips, id, every);
  long out = addrOf<HOME>(c, id, level);
  coarse.rhs[out] = St::st(rhs);
  float first = 0.f;
  if (every) {
    int e = coarse.eflag[out];
    if (!(e & STORED_SOLID)) first = mixv(0.f, (0.f - rhs) / fmaxv(diagonal(m.weight, e & 63), 1e-12f), 0.8f);
  }
  coarse.p[out] = St::st(first);
}
template <class St>
__device__ __forceinline__ void mg_restrict_bricks_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int level = code & 255;
  const int sd = m.lv[level].side, n = sd * sd * sd;
  MG_BRICK_GROUP(n)
  MG_GROUP_HOME
  (void)rec;
  for (int l = gr.first; l < n; l += gr.step) {
    int p[3];
    levelCell(c.home.bx, c.home.by, c.home.bz, sd, l, p);
    restrictCell<St, true>(c, m, level, code >> 8, p);
  }
}
FP_MG_PAIR(mg_restrict_bricks, 256)
template <class St>
__device__ __forceinline__ void mg_restrict_tiles_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int level = code & 255;
  const int sd = m.lv[level].side, n = sd * sd * sd;
  MG_TILE_GROUP(n)
  for (int l = gr.first; l < n; l += gr.step) {
    int p[3];
    levelCell(c.tc[0], c.tc[1], c.tc[2], sd, l, p);
    if (level - 1 >= m.L.brickLevels) restrictCell<St, true>(c, m, level, code >> 8, p);
    else restrictCell<St>(c, m, level, code >> 8, p);
  }
}
FP_MG_PAIR(mg_restrict_tiles, 64)

// coarseValue (pressure.wgsl:670-691): Neumann ghosts beside solids, a mirror below the floor.
// Each coarse cell is resolved once: its topology byte and value share the address.
// NEAR: p and its six neighbors lie in the context's neighborhood (a brick context at a
// brick level); reads the floor moved use the general lookup.
template <class St, bool NEAR = false>
__device__ __forceinline__ float coarseValue(const MGCtx &c, const MGArgsT<typename St::T> &m, const LevelView<typename St::T> &cv,
                                             const int p[3], int level) {
  const int groundY = m.P.ground ? 0 : (int)0x80000000;
  int cc[3] = {p[0], p[1] > groundY ? p[1] : groundY, p[2]};
  const long a0 = NEAR && cc[1] == p[1] ? addrOf<true>(c, cc, level) : addrAt(c, cc, level);
  if (!(a0 >= 0 && (cv.topo[a0] & STORED_SOLID))) return a0 < 0 ? 0.f : St::ld(cv.p[a0]);
  float sum = 0.f, count = 0.f;
  for (int axis = 0; axis < 3; axis++)
    for (int sd = -1; sd <= 1; sd += 2) {
      int q[3] = {cc[0], cc[1], cc[2]};
      q[axis] += sd;
      const int y = q[1];
      q[1] = q[1] > groundY ? q[1] : groundY;
      const long a = NEAR && cc[1] == p[1] && q[1] == y ? addrOf<true>(c, q, level) : addrAt(c, q, level);
      if (!(a >= 0 && (cv.topo[a] & STORED_SOLID))) {
        sum += a < 0 ? 0.f : St::ld(cv.p[a]);
        count += 1.0f;
      }
    }
  return sum / fmaxv(count, 1.0f);
}

// The coarse value a prolongation weighs at coarse cell cc: coarseValue, which without
// colliders is the value at cc with its y raised to the floor (no cell at or above the floor
// is solid then), so that read alone.
// NEAR: cc lies within one cell of the context's cube, read through its neighborhood (not
// for a cell the floor moves, nor for the solid rule's neighbors, which may lie farther).
// NEAR_SOLID: the solid rule's neighbors lie in the neighborhood too (coarseValue NEAR).
template <class St, bool NEAR = false, bool NEAR_SOLID = false>
__device__ __forceinline__ float prolongValue(const MGCtx &c, const MGArgsT<typename St::T> &m,
                                              const LevelView<typename St::T> &cv, const int cc[3], int level) {
  if (m.P.colliders != 0) return coarseValue<St, NEAR_SOLID>(c, m, cv, cc, level);
  const bool floor = m.P.ground && cc[1] < 0;
  if (NEAR && !floor) return readAt<St, true>(c, cv.p, cc, level);
  const int q[3] = {cc[0], floor ? 0 : cc[1], cc[2]};
  return readAt<St>(c, cv.p, q, level);
}

// prolongateCell (pressure.wgsl:726-755): trilinear correction from the coarse level, p to
// q. The fine cells of an item (a brick's or a tile's cube of `sd` cells a side) weigh the
// coarse cells from o - 1 to o + sd / 2 on each axis, o = its first fine cell / 2: those
// (sd / 2 + 2)^3 values are resolved once per item into LDS (`vals`, x fastest), each by
// one lane, before any fine cell interpolates them.
constexpr int PROLONG_VALUES = MG_GROUP * 27;  // the largest group: 32 items of 2^3 cells
template <class St>
__device__ __forceinline__ void stageProlong(const MGCtx &c, const MGArgsT<typename St::T> &m, int level, int flips,
                                             const int org[3], int sd, int first, int step, float *vals) {
  const auto cv = levelView(m, level + 1, flips);
  const int rs = sd / 2 + 2, rn = rs * rs * rs;
  for (int r = first; r < rn; r += step) {
    const int t = r / rs;
    const int cc[3] = {org[0] * (sd / 2) - 1 + (r - t * rs), org[1] * (sd / 2) - 1 + t % rs,
                       org[2] * (sd / 2) - 1 + t / rs};
    // The coarse cells lie within one cell of the item's coarse cube: in the home brick's
    // neighborhood at a brick level, the context tile's at a tile level.
    // At a brick level with at least two cells a brick, the solid rule's neighbors (one
    // cell farther) stay in the home neighborhood as well.
    if (level + 1 < m.L.brickLevels && c.hasBrick)
      vals[r] = prolongValue<St, true, true>(c, m, cv, cc, level + 1);
    else if (level + 1 >= m.L.brickLevels && c.hasTile)
      vals[r] = prolongValue<St, true>(c, m, cv, cc, level + 1);
    else
      vals[r] = prolongValue<St>(c, m, cv, cc, level + 1);
  }
}
template <class St>
__device__ __forceinline__ void prolongCell(const MGArgsT<typename St::T> &m, int level, int flips, const int org[3],
                                            int sd, const float *vals, const int p[3], size_t a) {
  const auto lv = levelView(m, level, flips);
  int e = lv.eflag[a];
  const float own = St::ld(lv.p[a]);
  if ((e & STORED_SOLID) || !(e & UNKNOWN)) {
    lv.q[a] = St::st(0.f);
    return;
  }
  const int rs = sd / 2 + 2;
  float f[3];
  int i[3];
  for (int k = 0; k < 3; k++) {
    float q = ((float)p[k] + 0.5f) * 0.5f - 0.5f;
    float fl = floorf(q);
    i[k] = (int)fl - (org[k] * (sd / 2) - 1);
    f[k] = q - fl;
  }
  const float *b = vals + (i[2] * rs + i[1]) * rs + i[0];
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = b[((k >> 2) * rs + ((k >> 1) & 1)) * rs + (k & 1)];
  float corr = mixv(mixv(mixv(v[0], v[1], f[0]), mixv(v[2], v[3], f[0]), f[1]),
                    mixv(mixv(v[4], v[5], f[0]), mixv(v[6], v[7], f[0]), f[1]), f[2]);
  lv.q[a] = St::st(own + corr);
}
template <class St>
__device__ __forceinline__ void mg_prolong_bricks_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int level = code & 255, flips = code >> 8;
  const int sd = m.lv[level].side, n = sd * sd * sd;
  MG_BRICK_GROUP(n)
  __shared__ int recs_[MG_GROUP][HOME_WORDS];
  __shared__ float vals[PROLONG_VALUES];
  loadGroupHomes(m.dir, m.active, count_, gr, recs_, m.L.side);
  const int rn = (sd / 2 + 2) * (sd / 2 + 2) * (sd / 2 + 2);
  float *mine_vals = vals + (gr.j < 0 ? 0 : gr.j) * rn;
  MGCtx c;
  c.dir = m.dir;
  c.L = m.L;
  c.side = m.L.side;
  c.hasBrick = true;
  int org[3] = {0, 0, 0}, slot = -1;
  if (mine) {
    const int id = m.active[bi];
    brickTile(c, id);
    c.home = homeOf(m.dir, id, recs_[gr.j]);
    org[0] = c.home.bx;
    org[1] = c.home.by;
    org[2] = c.home.bz;
    slot = entrySlot(recs_[gr.j][13], c.side);
    stageProlong<St>(c, m, level, flips, org, sd, gr.first, gr.step, mine_vals);
  }
  __syncthreads();
  if (!mine) return;
  for (int l = gr.first; l < n; l += gr.step) {
    int p[3];
    levelCell(org[0], org[1], org[2], sd, l, p);
    prolongCell<St>(m, level, flips, org, sd, mine_vals, p, (size_t)slot * n + l);
  }
}
FP_MG_PAIR(mg_prolong_bricks, 256)
template <class St>
__device__ __forceinline__ void mg_prolong_tiles_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int level = code & 255, flips = code >> 8;
  const int sd = m.lv[level].side, n = sd * sd * sd;
  const int k_ = groupSize(n, (int)blockDim.x);
  if (vb * k_ >= m.dir.capacity) return;
  const Group gr = groupOf(vb, n);
  const int t = gr.g * gr.k + gr.j;
  __shared__ int tnbs_[MG_GROUP][27];
  __shared__ float vals[PROLONG_VALUES];
  for (int e = threadIdx.x; e < gr.k * 27; e += blockDim.x) {
    const int jj = e / 27, w = e - jj * 27, tt = gr.g * gr.k + jj;
    if (tt < m.dir.capacity) tnbs_[jj][w] = m.dir.tileRec[(size_t)tt * TILE_RECORD_WORDS + 4 + w];
  }
  __syncthreads();
  const int rn = (sd / 2 + 2) * (sd / 2 + 2) * (sd / 2 + 2);
  float *mine_vals = vals + (gr.j < 0 ? 0 : gr.j) * rn;
  const int *trec = m.dir.tileRec + (size_t)(t < m.dir.capacity ? t : 0) * TILE_RECORD_WORDS;
  const bool live = gr.j >= 0 && t < m.dir.capacity && trec[3] != 0;
  MGCtx c;
  c.dir = m.dir;
  c.L = m.L;
  c.side = m.L.side;
  c.hasBrick = false;
  c.hasTile = true;
  c.tc[0] = trec[0];
  c.tc[1] = trec[1];
  c.tc[2] = trec[2];
  c.tnb = tnbs_[gr.j < 0 ? 0 : gr.j];
  if (live) stageProlong<St>(c, m, level, flips, c.tc, sd, gr.first, gr.step, mine_vals);
  __syncthreads();
  if (!live) return;
  for (int l = gr.first; l < n; l += gr.step) {
    int p[3];
    levelCell(c.tc[0], c.tc[1], c.tc[2], sd, l, p);
    prolongCell<St>(m, level, flips, c.tc, sd, mine_vals, p, (size_t)t * n + l);
  }
}
FP_MG_PAIR(mg_prolong_tiles, 64)

// relaxCoarse (pressure.wgsl:549-599): every bottom iteration in one workgroup, up to
// BOTTOM64_TILES / 1024 tile slots per lane (lane, lane + 1024, ...), values in LDS. The
// result lands in the bottom level's current iterate. Sweeps alternate between two LDS
// copies (read one, write the other), so one barrier separates them.
constexpr int BOTTOM_LANES = 1024;
// PER: the tile slots a lane holds in registers (the capacity over 1024, rounded up to a
// power of two by the caller), so that a lane runs only the rounds the capacity needs.
template <class St, int PER>
__device__ __forceinline__ void bottomSweeps(const MGArgsT<typename St::T> &m, int code, float (*values)[BOTTOM64_TILES]) {
  float *value = values[0];
  const auto lv = levelView(m, m.bottom, code >> 8);
  const int cap = m.dir.capacity;
  const int per = (cap + BOTTOM_LANES - 1) / BOTTOM_LANES;
  const float w = m.weight;
  float mine[PER], rhs[PER];
  int flags[PER], nbs[PER][6];
  bool valid[PER], live[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int t = threadIdx.x + k * BOTTOM_LANES;
    const bool in = k < per && t < cap;
    const int *trec = m.dir.tileRec + (size_t)(in ? t : 0) * TILE_RECORD_WORDS;
    valid[k] = in && trec[3] != 0;
    const int e = valid[k] ? lv.eflag[t] : 0;
    live[k] = valid[k] && (e & UNKNOWN) && !(e & STORED_SOLID);
    flags[k] = e & 63;
    rhs[k] = valid[k] ? St::ld(lv.rhs[t]) : 0.f;
    mine[k] = valid[k] && (e & UNKNOWN) ? St::ld(lv.p[t]) : 0.f;
    if (in) value[t] = mine[k];
    // The open sides' neighbor slots, read once (-1: closed or no neighbor).
#pragma unroll
    for (int sd = 0; sd < 6; sd++)
      nbs[k][sd] = live[k] && ((flags[k] >> sd) & 1)
                       ? trec[4 + (OFF[sd][2] + 1) * 9 + (OFF[sd][1] + 1) * 3 + (OFF[sd][0] + 1)]
                       : -1;
  }
  __syncthreads();
  for (int it = 0; it < BOTTOM_ITERATIONS; it++) {
    const float *cur = values[it & 1];
    float *next = values[(it & 1) ^ 1];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      float out = 0.f;
      if (live[k]) {
        float n[6];
        for (int sd = 0; sd < 6; sd++) n[sd] = nbs[k][sd] >= 0 ? cur[nbs[k][sd]] : 0.f;
        float sum = w * (n[0] + n[1]) + w * (n[2] + n[3]) + w * (n[4] + n[5]);
        float cand = (sum - rhs[k]) / fmaxv(diagonal(w, flags[k]), 1e-12f);
        out = mixv(mine[k], cand, 0.8f);
      }
      const int t = threadIdx.x + k * BOTTOM_LANES;
      mine[k] = out;
      if (k < per && t < cap) next[t] = out;
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < PER; k++)
    if (valid[k]) lv.p[threadIdx.x + k * BOTTOM_LANES] = St::st(mine[k]);
}
template <class St> __device__ __forceinline__ void mg_bottom_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  if (vb != 0 || m.tilesFused) return;
  __shared__ float values[2][BOTTOM64_TILES];
  const int per = (m.dir.capacity + BOTTOM_LANES - 1) / BOTTOM_LANES;
  if (per <= 1) bottomSweeps<St, 1>(m, code, values);
  else if (per <= 2) bottomSweeps<St, 2>(m, code, values);
  else if (per <= 4) bottomSweeps<St, 4>(m, code, values);
  else bottomSweeps<St, BOTTOM64_TILES / BOTTOM_LANES>(m, code, values);
}
FP_MG_PAIR(mg_bottom, 1024)

// The coarsest level's sweeps in FP64: the same weighted Jacobi as relaxCoarse, up to
// BOTTOM64_TILES / 1024 tile slots per lane, values in LDS. The result is rounded to FP32,
// then stored.
__device__ __forceinline__ double mix64(double a, double b, double t) { return a * (1.0 - t) + b * t; }
template <class St, int PER>
__device__ __forceinline__ void bottom64Sweeps(const MGArgsT<typename St::T> &m, int code, double *value) {
  const auto lv = levelView(m, m.bottom, code >> 8);
  const int cap = m.dir.capacity;
  const double w = (double)m.weight;
  double mine[PER], rhs[PER];
  int flags[PER];
  bool live[PER];
  // Lanes take tile slots lane, lane + 1024, ...; `per` rounds cover the capacity.
  const int per = (cap + BOTTOM_LANES - 1) / BOTTOM_LANES;
  for (int k = 0; k < PER; k++) {
    const int t = threadIdx.x + k * BOTTOM_LANES;
    bool valid = k < per && t < cap && m.dir.tileRec[(size_t)t * TILE_RECORD_WORDS + 3] != 0;
    int e = valid ? lv.eflag[t] : 0;
    live[k] = valid && (e & UNKNOWN) && !(e & STORED_SOLID);
    flags[k] = e & 63;
    rhs[k] = valid ? (double)St::ld(lv.rhs[t]) : 0.0;
    mine[k] = valid && (e & UNKNOWN) ? (double)St::ld(lv.p[t]) : 0.0;
    if (k < per && t < cap) value[t] = mine[k];
  }
  __syncthreads();
  for (int it = 0; it < BOTTOM_ITERATIONS; it++) {
    double next[PER];
    for (int k = 0; k < PER; k++) {
      const int t = threadIdx.x + k * BOTTOM_LANES;
      next[k] = 0.0;
      if (!live[k]) continue;
      const int *trec = m.dir.tileRec + (size_t)t * TILE_RECORD_WORDS;
      double n[6];
      for (int sd = 0; sd < 6; sd++) {
        n[sd] = 0.0;
        if (!((flags[k] >> sd) & 1)) continue;
        int nb = trec[4 + (OFF[sd][2] + 1) * 9 + (OFF[sd][1] + 1) * 3 + (OFF[sd][0] + 1)];
        if (nb >= 0) n[sd] = value[nb];
      }
      double sum = w * (n[0] + n[1]) + w * (n[2] + n[3]) + w * (n[4] + n[5]);
      double diag = w * (double)openCount(flags[k], 0) + w * (double)openCount(flags[k], 2) +
                    w * (double)openCount(flags[k], 4);
      double cand = (sum - rhs[k]) / (diag < 1e-12 ? 1e-12 : diag);
      next[k] = mix64(mine[k], cand, 0.8);
    }
    __syncthreads();
    for (int k = 0; k < PER; k++) {
      const int t = threadIdx.x + k * BOTTOM_LANES;
      mine[k] = next[k];
      if (k < per && t < cap) value[t] = next[k];
    }
    __syncthreads();
  }
  for (int k = 0; k < PER; k++) {
    const int t = threadIdx.x + k * BOTTOM_LANES;
    if (k < per && t < cap && m.dir.tileRec[(size_t)t * TILE_RECORD_WORDS + 3] != 0) lv.p[t] = St::st((float)mine[k]);
  }
}
template <class St>
__device__ __forceinline__ void mg_bottom64_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  if (vb != 0 || m.tilesFused) return;
  __shared__ double value[BOTTOM64_TILES];
  const int per = (m.dir.capacity + BOTTOM_LANES - 1) / BOTTOM_LANES;
  if (per <= 1) bottom64Sweeps<St, 1>(m, code, value);
  else if (per <= 2) bottom64Sweeps<St, 2>(m, code, value);
  else if (per <= 4) bottom64Sweeps<St, 4>(m, code, value);
  else bottom64Sweeps<St, BOTTOM64_TILES / BOTTOM_LANES>(m, code, value);
}
FP_MG_PAIR(mg_bottom64, 1024)

// ---- the first tile level, a tile and its halo per workgroup -----------------------------
// The first tile level (one cell per brick, 4^3 per tile slot) runs its two sweeps and the
// restriction below it in one launch (k_mg_tiles_down), and the prolongation into it and its
// three sweeps in another (k_mg_tiles_up), instead of one launch per stage. A workgroup takes
// one live tile slot and the level's cells within three of it (10^3 cells from the tile's
// 27-neighborhood) into LDS; every sweep updates one cell less on each side, the way
// successive sweeps of the per-level kernels would, so the tile's own cells end exactly as
// they would there (the halo cells are computed again by the neighboring tiles' workgroups).
// Cells of tiles that are not live keep their values, as the per-level kernels never write
// them; cells of no tile read as zero. Each kernel stores the tile's cells into the level's
// other array (k_mg_tiles_down after the two sweeps, k_mg_tiles_up at the end), since the
// workgroups of neighboring tiles still read the current one: the caller flips the level.
constexpr int RG = 10, RG3 = RG * RG * RG;
// A value as its level's storage reads it back.
template <class St> __device__ __forceinline__ float stored(float x) { return St::ld(St::st(x)); }
struct TileRegion {
  static constexpr int N = RG;
  float v[2][RG3];
  float rhs[RG3];
  uint8_t flag[RG3];
  int8_t state[RG3];     // 1: a cell of a live tile, 0: of a tile that is not live, -1: of no tile
  int nb[27], live[27];  // the tile's 27 neighbor slots and whether each is live
  float coarse[216];     // k_mg_tiles_up: the 6^3 coarse values its prolongation weighs
};
// The region's cells from the 27 owners in R.nb (slots of a level whose cubes are 4^3, -1
// for none) and R.live, after a barrier that published them. Four cells a lane per pass,
// their loads issued before the first LDS store.
template <class St>
__device__ __forceinline__ void fillRegion(const LevelView<typename St::T> &lv, TileRegion &R) {
  constexpr int U = 4;
  const int lanes = (int)blockDim.x;
  for (int first = (int)threadIdx.x; first < RG3; first += U * lanes) {
    float pv[U], rv[U];
    int fv[U], sv[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int i = first + j * lanes;
      pv[j] = rv[j] = 0.f;
      fv[j] = 0;
      sv[j] = -1;
      if (i < RG3) {
        const int r[3] = {i % RG, (i / RG) % RG, i / (RG * RG)};
        int q = 0, local = 0;
        for (int a = 2; a >= 0; a--) {
          const int g = r[a] + 1;  // cell origin * 4 - 3 + r: owner offset (g >> 2) - 1, local g & 3
          q = q * 3 + (g >> 2);
          local = local * 4 + (g & 3);
        }
        const int nt = R.nb[q];
        if (nt >= 0) {
          const size_t a = (size_t)nt * 64 + local;
          pv[j] = St::ld(lv.p[a]);
          rv[j] = St::ld(lv.rhs[a]);
          fv[j] = lv.eflag[a];
          sv[j] = R.live[q] ? 1 : 0;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int i = first + j * lanes;
      if (i < RG3) {
        R.v[0][i] = pv[j];
        R.rhs[i] = rv[j];
        R.flag[i] = (uint8_t)fv[j];
        R.state[i] = (int8_t)sv[j];
      }
    }
  }
  __syncthreads();
}
// Loads the region of tile slot t at level L (current iterate `cur`) into LDS; returns false
// (for the whole workgroup) when t is not a live tile slot.
template <class St>
__device__ __forceinline__ bool loadRegion(const MGArgsT<typename St::T> &m, const LevelView<typename St::T> &lv,
                                           int t, TileRegion &R, int tc[3]) {
  const int *trec = m.dir.tileRec + (size_t)t * TILE_RECORD_WORDS;
  if (trec[3] == 0) return false;
  tc[0] = trec[0];
  tc[1] = trec[1];
  tc[2] = trec[2];
  if (threadIdx.x < 27) {
    const int nt = trec[4 + threadIdx.x];
    R.nb[threadIdx.x] = nt;
    R.live[threadIdx.x] = nt >= 0 && m.dir.tileRec[(size_t)nt * TILE_RECORD_WORDS + 3] != 0;
  }
  __syncthreads();
  fillRegion<St>(lv, R);
  return true;
}
// Whether region cell i keeps its value through a sweep (a cell of a tile that is not live, or
// of no tile).
__device__ __forceinline__ bool held(const TileRegion &R, int i) { return R.state[i] != 1; }
template <int SD> struct BrickRegion;
template <int SD> __device__ __forceinline__ bool held(const BrickRegion<SD> &R, int i);
// One weighted Jacobi sweep (relaxCell) over region cells [lo, N - lo)^3 of a region N
// cells a side, src to dst.
template <class St, class Rg>
__device__ __forceinline__ void regionSweep(Rg &R, int src, int lo, float w) {
  constexpr int RG = Rg::N;
  const int n = RG - 2 * lo, n3 = n * n * n;
  const float *a = R.v[src];
  float *d = R.v[src ^ 1];
  for (int j = (int)threadIdx.x; j < n3; j += (int)blockDim.x) {
    const int i = (lo + j % n) + RG * (lo + (j / n) % n) + RG * RG * (lo + j / (n * n));
    if (held(R, i)) {
      d[i] = a[i];
      continue;
    }
    const int e = R.flag[i];
    if ((e & STORED_SOLID) || !(e & UNKNOWN)) {
      d[i] = 0.f;
      continue;
    }
    const float nb[6] = {a[i + 1], a[i - 1], a[i + RG], a[i - RG], a[i + RG * RG], a[i - RG * RG]};
    const float cand = (openSum(nb, e & 63, w) - R.rhs[i]) / fmaxv(diagonal(w, e & 63), 1e-12f);
    d[i] = stored<St>(mixv(a[i], cand, 0.8f));
  }
  __syncthreads();
}
// The tile's 64 cells (region [3, 7)^3) from R.v[src] into lv.q.
template <class St>
__device__ __forceinline__ void storeTile(const TileRegion &R, int src, const LevelView<typename St::T> &lv, int t) {
  if (threadIdx.x < 64) {
    const int l = (int)threadIdx.x;
    const int i = (3 + (l & 3)) + RG * (3 + ((l >> 2) & 3)) + RG * RG * (3 + (l >> 4));
    lv.q[(size_t)t * 64 + l] = St::st(R.v[src][i]);
  }
}
// prolongateCell (pressure.wgsl:726-755) over the whole region of the cube at `org` (in
// cubes), R.v[0] to R.v[1], from the 6^3 coarse values in R.coarse.
template <class St>
__device__ __forceinline__ void regionProlong(TileRegion &R, const int org[3]) {
  for (int i = (int)threadIdx.x; i < RG3; i += (int)blockDim.x) {
    if (R.state[i] != 1) {
      R.v[1][i] = R.v[0][i];
      continue;
    }
    const int e = R.flag[i];
    if ((e & STORED_SOLID) || !(e & UNKNOWN)) {
      R.v[1][i] = 0.f;
      continue;
    }
    const int p[3] = {org[0] * 4 - 3 + i % RG, org[1] * 4 - 3 + (i / RG) % RG, org[2] * 4 - 3 + i / (RG * RG)};
    float f[3];
    int b[3];
    for (int k = 0; k < 3; k++) {
      const float q = ((float)p[k] + 0.5f) * 0.5f - 0.5f;
      const float fl = floorf(q);
      b[k] = (int)fl - (org[k] * 2 - 2);
      f[k] = q - fl;
    }
    const float *c0 = R.coarse + (b[2] * 6 + b[1]) * 6 + b[0];
    const float v[8] = {c0[0], c0[1], c0[6], c0[7], c0[36], c0[37], c0[42], c0[43]};
    const float corr = mixv(mixv(mixv(v[0], v[1], f[0]), mixv(v[2], v[3], f[0]), f[1]),
                            mixv(mixv(v[4], v[5], f[0]), mixv(v[6], v[7], f[0]), f[1]), f[2]);
    R.v[1][i] = stored<St>(R.v[0][i] + corr);
  }
  __syncthreads();
}
// code: the flips (bits 0-15).
template <class St>
__device__ __forceinline__ void mg_tiles_down_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int Lt = m.L.brickLevels, t = vb;
  if (t >= m.dir.capacity) return;
  __shared__ TileRegion R;
  const auto lv = levelView(m, Lt, code);
  int tc[3];
  if (!loadRegion<St>(m, lv, t, R, tc)) return;
  const float w = m.weight;
  regionSweep<St>(R, 0, 1, w);
  regionSweep<St>(R, 1, 2, w);
  storeTile<St>(R, 0, lv, t);
  // restrictCell into the tile's 2^3 cells of the next level, from the swept values.
  if (threadIdx.x < 8) {
    const auto coarse = levelView(m, Lt + 1, code);
    const int l = (int)threadIdx.x, cx = l & 1, cy = (l >> 1) & 1, cz = l >> 2;
    const float *a = R.v[0];
    const int i0 = (3 + 2 * cx) + RG * (3 + 2 * cy) + RG * RG * (3 + 2 * cz);
    float residual = 0.f;
    bool every = true;
    for (int k = 0; k < 8; k++) {
      const int i = i0 + (k & 1) + RG * ((k >> 1) & 1) + RG * RG * (k >> 2);
      const int e = R.flag[i];
      const bool unknown = (e & UNKNOWN) != 0;
      every = every && unknown;
      if (!unknown || (e & STORED_SOLID)) continue;
      const int flags = e & 63;
      const float nb[6] = {a[i + 1], a[i - 1], a[i + RG], a[i - RG], a[i + RG * RG], a[i - RG * RG]};
      residual += R.rhs[i] - (openSum(nb, flags, w) - diagonal(w, flags) * a[i]);
    }
    const float rhs = residual * 0.5f;
    const size_t out = (size_t)t * 8 + l;
    coarse.rhs[out] = St::st(rhs);
    float first = 0.f;
    if (every) {
      const int e = coarse.eflag[out];
      if (!(e & STORED_SOLID)) first = mixv(0.f, (0.f - rhs) / fmaxv(diagonal(w, e & 63), 1e-12f), 0.8f);
    }
    coarse.p[out] = St::st(first);
  }
}
FP_MG_PAIR(mg_tiles_down, 256)
template <class St>
__device__ __forceinline__ void mg_tiles_up_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int Lt = m.L.brickLevels, t = vb;
  if (t >= m.dir.capacity) return;
  __shared__ TileRegion R;
  const auto lv = levelView(m, Lt, code);
  int tc[3];
  if (!loadRegion<St>(m, lv, t, R, tc)) return;
  // The coarse cells [tc * 2 - 2, tc * 2 + 4) on each axis, resolved once (prolongValue).
  {
    MGCtx c;
    c.dir = m.dir;
    c.L = m.L;
    c.side = m.L.side;
    c.hasBrick = false;
    c.hasTile = true;
    c.tc[0] = tc[0];
    c.tc[1] = tc[1];
    c.tc[2] = tc[2];
    c.tnb = R.nb;
    const auto cv = levelView(m, Lt + 1, code);
    for (int j = (int)threadIdx.x; j < 216; j += (int)blockDim.x) {
      const int cc[3] = {tc[0] * 2 - 2 + j % 6, tc[1] * 2 - 2 + (j / 6) % 6, tc[2] * 2 - 2 + j / 36};
      R.coarse[j] = prolongValue<St, true>(c, m, cv, cc, Lt + 1);
    }
  }
  __syncthreads();
  regionProlong<St>(R, tc);
  const float w = m.weight;
  regionSweep<St>(R, 1, 1, w);
  regionSweep<St>(R, 0, 2, w);
  regionSweep<St>(R, 1, 3, w);
  storeTile<St>(R, 0, lv, t);
}
FP_MG_PAIR(mg_tiles_up, 256)

// ---- brick levels by tile of bricks and halo ---------------------------------------------
// A brick level whose cubes hold SD^3 cells runs its sweeps with the restriction below them
// (k_mg_bricks_down) and the prolongation into it with its three sweeps (k_mg_bricks_up) in
// one launch each, as the first tile level's halo kernels do. They run at the last brick
// level when it holds 2^3 cells a brick (above a first tile level of one cell a brick). At
// level 0 of 4^3 cells (SD = 4: 22^3 cells and 146 KB of LDS a tile, one workgroup a CU) the
// same scheme took 73 us against 65 us for the launches it would replace: the regions of
// sparse tiles are mostly empty cells, which every sweep still visits. A workgroup takes one tile slot in use: the level's
// (4 SD)^3 cells of its 4^3 bricks and the cells within three of them, whose pool slots
// come from the 27 neighbor tiles' brick entries. Every brick with a pool slot is active and
// swept by the per-level kernels, so cells of a brick with a slot are updated, and cells of
// a brick without one read as zero: their flag byte is zero (not an unknown), so a sweep or
// the prolongation writes them zero again, and no per-cell state is kept. Each kernel stores
// the tile's bricks' cells into the level's other array; the caller flips the level.
template <int SD> struct BrickRegion {
  static constexpr int N = 4 * SD + 6, N3 = N * N * N;
  static constexpr int B0 = SD == 2 ? -2 : -1, NB = SD == 2 ? 8 : 6;  // region bricks B0 .. B0 + NB - 1
  static constexpr int NC = SD == 2 ? 8 : 12;  // coarse cells a side the prolongation weighs
  float v[2][N3];
  float rhs[N3];
  uint8_t flag[N3];
  int slot[NB * NB * NB];  // the region's bricks (tile-relative B0 .. B0 + NB - 1 on each axis)
  int tnb[27];
  float coarse[NC * NC * NC];  // the coarse values the prolongation weighs
};
template <int SD> __device__ __forceinline__ bool held(const BrickRegion<SD> &, int) { return false; }
// The region's pool slot of the brick holding tile-relative cell g (tile-relative bricks
// B0 ..), -1 without one, and the cell's index inside its brick.
template <int SD>
__device__ __forceinline__ int regionSlot(const BrickRegion<SD> &R, const int g[3], int &local) {
  constexpr int sh = SD == 2 ? 1 : 2, NB = BrickRegion<SD>::NB, B0 = BrickRegion<SD>::B0;
  local = (((g[2] & (SD - 1)) * SD + (g[1] & (SD - 1))) * SD) + (g[0] & (SD - 1));
  return R.slot[(((g[2] >> sh) - B0) * NB + ((g[1] >> sh) - B0)) * NB + ((g[0] >> sh) - B0)];
}
// Loads level L's region around tile slot t; false (for the whole workgroup) when t is not
// in use. tc: the tile's coordinates.
template <class St, int SD>
__device__ __forceinline__ bool loadBrickRegion(const MGArgsT<typename St::T> &m, const LevelView<typename St::T> &lv,
                                                int t, BrickRegion<SD> &R, int tc[3]) {
  using Rg = BrickRegion<SD>;
  const int *trec = m.dir.tileRec + (size_t)t * TILE_RECORD_WORDS;
  if (trec[3] == 0) return false;
  tc[0] = trec[0];
  tc[1] = trec[1];
  tc[2] = trec[2];
  if (threadIdx.x < 27) R.tnb[threadIdx.x] = trec[4 + threadIdx.x];
  __syncthreads();
  for (int j = (int)threadIdx.x; j < Rg::NB * Rg::NB * Rg::NB; j += (int)blockDim.x) {
    // Tile-relative brick b: tile offset b >> 2, brick b & 3.
    const int b[3] = {j % Rg::NB + Rg::B0, (j / Rg::NB) % Rg::NB + Rg::B0, j / (Rg::NB * Rg::NB) + Rg::B0};
    const int nt = R.tnb[((b[2] >> 2) + 1) * 9 + ((b[1] >> 2) + 1) * 3 + ((b[0] >> 2) + 1)];
    R.slot[j] = nt < 0 ? -1 : entrySlot(m.dir.entries[(size_t)nt * 64 + (((b[2] & 3) * 4 + (b[1] & 3)) * 4 + (b[0] & 3))],
                                        m.L.side);
  }
  __syncthreads();
  // Every cell's loads in one pass: the region's LDS, not registers, bounds the resident
  // workgroups.
  constexpr int N = Rg::N, N3 = Rg::N3, n = SD * SD * SD;
  const int lanes = (int)blockDim.x;
  constexpr int U = SD == 2 ? (N3 + 255) / 256 : (N3 + 1023) / 1024;
  for (int first = (int)threadIdx.x; first < N3; first += U * lanes) {
    float pv[U], rv[U];
    int fv[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int i = first + j * lanes;
      pv[j] = rv[j] = 0.f;
      fv[j] = 0;
      if (i < N3) {
        const int g[3] = {i % N - 3, (i / N) % N - 3, i / (N * N) - 3};  // tile-relative cell
        int local;
        const int sl = regionSlot(R, g, local);
        if (sl >= 0) {
          const size_t a = (size_t)sl * n + local;
          pv[j] = St::ld(lv.p[a]);
          rv[j] = St::ld(lv.rhs[a]);
          fv[j] = lv.eflag[a];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int i = first + j * lanes;
      if (i < N3) {
        R.v[0][i] = pv[j];
        R.rhs[i] = rv[j];
        R.flag[i] = (uint8_t)fv[j];
      }
    }
  }
  __syncthreads();
  return true;
}
// The tile's bricks' cells (region [3, 3 + 4 SD)^3) from R.v[src] into lv.q.
template <class St, int SD>
__device__ __forceinline__ void storeBricks(const BrickRegion<SD> &R, int src, const LevelView<typename St::T> &lv) {
  constexpr int N = BrickRegion<SD>::N, T = 4 * SD, n = SD * SD * SD;
  for (int l = (int)threadIdx.x; l < T * T * T; l += (int)blockDim.x) {
    const int g[3] = {l % T, (l / T) % T, l / (T * T)};
    int local;
    const int sl = regionSlot(R, g, local);
    if (sl < 0) continue;
    const int i = (3 + g[0]) + N * (3 + g[1]) + N * N * (3 + g[2]);
    lv.q[(size_t)sl * n + local] = St::st(R.v[src][i]);
  }
}
// prolongateCell over the whole region into R.v[1], from the NC^3 coarse values in R.coarse
// (coarse cells from the tile's coarse origin - 2; the coarse level holds 2 SD cells a tile
// side).
template <class St, int SD>
__device__ __forceinline__ void brickRegionProlong(BrickRegion<SD> &R, const int tc[3]) {
  constexpr int N = BrickRegion<SD>::N, N3 = BrickRegion<SD>::N3, NC = BrickRegion<SD>::NC;
  for (int i = (int)threadIdx.x; i < N3; i += (int)blockDim.x) {
    const int e = R.flag[i];
    if ((e & STORED_SOLID) || !(e & UNKNOWN)) {
      R.v[1][i] = 0.f;
      continue;
    }
    const int p[3] = {tc[0] * 4 * SD - 3 + i % N, tc[1] * 4 * SD - 3 + (i / N) % N, tc[2] * 4 * SD - 3 + i / (N * N)};
    float f[3];
    int b[3];
    for (int k = 0; k < 3; k++) {
      const float q = ((float)p[k] + 0.5f) * 0.5f - 0.5f;
      const float fl = floorf(q);
      b[k] = (int)fl - (tc[k] * 2 * SD - 2);
      f[k] = q - fl;
    }
    const float *c0 = R.coarse + (b[2] * NC + b[1]) * NC + b[0];
    const float v[8] = {c0[0], c0[1], c0[NC], c0[NC + 1], c0[NC * NC], c0[NC * NC + 1], c0[NC * NC + NC],
                        c0[NC * NC + NC + 1]};
    const float corr = mixv(mixv(mixv(v[0], v[1], f[0]), mixv(v[2], v[3], f[0]), f[1]),
                            mixv(mixv(v[4], v[5], f[0]), mixv(v[6], v[7], f[0]), f[1]), f[2]);
    R.v[1][i] = stored<St>(R.v[0][i] + corr);
  }
  __syncthreads();
}
// The coarse values the prolongation into level L weighs (prolongValue at level L + 1).
template <class St, int SD>
__device__ __forceinline__ void stageBrickCoarse(const MGArgsT<typename St::T> &m, int L, int code, BrickRegion<SD> &R,
                                                 const int tc[3]) {
  constexpr int NC = BrickRegion<SD>::NC;
  MGCtx c;
  c.dir = m.dir;
  c.L = m.L;
  c.side = m.L.side;
  c.hasBrick = false;
  c.hasTile = true;
  c.tc[0] = tc[0];
  c.tc[1] = tc[1];
  c.tc[2] = tc[2];
  c.tnb = R.tnb;
  const auto cv = levelView(m, L + 1, code);
  for (int j = (int)threadIdx.x; j < NC * NC * NC; j += (int)blockDim.x) {
    const int cc[3] = {tc[0] * 2 * SD - 2 + j % NC, tc[1] * 2 * SD - 2 + (j / NC) % NC, tc[2] * 2 * SD - 2 + j / (NC * NC)};
    R.coarse[j] = prolongValue<St, true>(c, m, cv, cc, L + 1);
  }
  __syncthreads();
}
// code: the flips (bits 0-15). Level L = brickLevels - 1, 2^3 cells a brick.
template <class St>
__device__ __forceinline__ void mg_bricks_down_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int L = m.L.brickLevels - 1, t = vb;
  if (t >= m.dir.capacity) return;
  __shared__ BrickRegion<2> R;
  constexpr int BR = BrickRegion<2>::N;
  const auto lv = levelView(m, L, code);
  int tc[3];
  if (!loadBrickRegion<St>(m, lv, t, R, tc)) return;
  const float w = m.weight;
  regionSweep<St>(R, 0, 1, w);
  regionSweep<St>(R, 1, 2, w);
  storeBricks<St>(R, 0, lv);
  // restrictCell into the first tile level's cell of each of the tile's 64 bricks, from the
  // swept values (its eight children are the brick's 2^3 cells).
  if (threadIdx.x < 64) {
    const auto coarse = levelView(m, L + 1, code);
    const int l = (int)threadIdx.x, bx = l & 3, by = (l >> 2) & 3, bz = l >> 4;
    const float *a = R.v[0];
    const int i0 = (3 + 2 * bx) + BR * (3 + 2 * by) + BR * BR * (3 + 2 * bz);
    float residual = 0.f;
    bool every = true;
    for (int k = 0; k < 8; k++) {
      const int i = i0 + (k & 1) + BR * ((k >> 1) & 1) + BR * BR * (k >> 2);
      const int e = R.flag[i];
      const bool unknown = (e & UNKNOWN) != 0;
      every = every && unknown;
      if (!unknown || (e & STORED_SOLID)) continue;
      const int flags = e & 63;
      const float nb[6] = {a[i + 1], a[i - 1], a[i + BR], a[i - BR], a[i + BR * BR], a[i - BR * BR]};
      residual += R.rhs[i] - (openSum(nb, flags, w) - diagonal(w, flags) * a[i]);
    }
    const float rhs = residual * 0.5f;
    const size_t out = (size_t)t * 64 + l;
    coarse.rhs[out] = St::st(rhs);
    float first = 0.f;
    if (every) {
      const int e = coarse.eflag[out];
      if (!(e & STORED_SOLID)) first = mixv(0.f, (0.f - rhs) / fmaxv(diagonal(w, e & 63), 1e-12f), 0.8f);
    }
    coarse.p[out] = St::st(first);
  }
}
FP_MG_PAIR(mg_bricks_down, 256)
// The prolongation into brick level L (SD^3 cells a brick) and its three sweeps.
template <class St, int SD>
__device__ __forceinline__ void bricksUp(const MGArgsT<typename St::T> &m, int L, int t, int code, BrickRegion<SD> &R) {
  const auto lv = levelView(m, L, code);
  int tc[3];
  if (!loadBrickRegion<St>(m, lv, t, R, tc)) return;
  stageBrickCoarse<St>(m, L, code, R, tc);
  brickRegionProlong<St>(R, tc);
  const float w = m.weight;
  regionSweep<St>(R, 1, 1, w);
  regionSweep<St>(R, 0, 2, w);
  regionSweep<St>(R, 1, 3, w);
  storeBricks<St>(R, 0, lv);
}
template <class St>
__device__ __forceinline__ void mg_bricks_up_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  if (vb >= m.dir.capacity) return;
  __shared__ BrickRegion<2> R;
  bricksUp<St>(m, m.L.brickLevels - 1, vb, code, R);
}
FP_MG_PAIR(mg_bricks_up, 256)


// ---- the coarser tile levels of a V-cycle in one workgroup -------------------------------
// For a simulation of at most TC_MAX tile slots, the V-cycle below the first tile level
// (PressureSolver.ts:303-363 from the second tile level down: two sweeps, the restriction,
// the bottom iterations, the prolongation, three sweeps) runs in one 1024-lane workgroup,
// both levels held in LDS. Each stage is the arithmetic of the per-level kernels, cell by
// cell, with their storage rounding: LDS keeps each value as its level's storage reads it
// back. Barriers separate the stages. The second tile level's final iterate is stored back
// to the array it was read from (the per-level kernels flip it six times), where
// k_mg_tiles_up reads it.
struct TileCycleLds {
  float v[2][TC_CELLS];  // iterates; level k's current one is v[cur[k]]
  float rhs[TC_CELLS];
  uint8_t flag[TC_CELLS];
  uint8_t topo[TC_CELLS];
  int tnb[TC_MAX * 27];  // each tile slot's 27 neighbor slots
  int tc[TC_MAX * 4];    // tile coordinates and the live word
  double b64[2][TC_MAX];
};
// Level k (1: the second tile level, 2 cells a side per tile slot; 2: the bottom, one).
__device__ __forceinline__ int tcOffset(int C, int k) { return k == 1 ? 0 : C * 8; }
// The LDS index of the level-k cell at level coordinates g (the home tile t's neighborhood,
// else the directory's hash), -1 for a cell of no tile (addrAt of the tile levels).
__device__ __forceinline__ int tcIndex(const TileCycleLds &S, const Dir &dir, int C, int k, int t, const int g[3]) {
  const int sh = 2 - k, s = 1 << sh;
  const int d[3] = {(g[0] >> sh) - S.tc[t * 4], (g[1] >> sh) - S.tc[t * 4 + 1], (g[2] >> sh) - S.tc[t * 4 + 2]};
  int nt;
  if ((unsigned)(d[0] + 1) <= 2u && (unsigned)(d[1] + 1) <= 2u && (unsigned)(d[2] + 1) <= 2u)
    nt = S.tnb[t * 27 + (d[2] + 1) * 9 + (d[1] + 1) * 3 + (d[0] + 1)];
  else
    nt = findTile(dir, g[0] >> sh, g[1] >> sh, g[2] >> sh);
  if (nt < 0) return -1;
  return tcOffset(C, k) + (nt << (3 * sh)) + (((g[2] & (s - 1)) << sh | (g[1] & (s - 1))) << sh) + (g[0] & (s - 1));
}
// The level coordinates of cell i of level k (tile slot t = i / cells).
__device__ __forceinline__ void tcCell(const TileCycleLds &S, int k, int i, int &t, int g[3]) {
  const int sh = 2 - k, l = i & ((1 << (3 * sh)) - 1), s = 1 << sh;
  t = i >> (3 * sh);
  g[0] = (S.tc[t * 4] << sh) + (l & (s - 1));
  g[1] = (S.tc[t * 4 + 1] << sh) + ((l >> sh) & (s - 1));
  g[2] = (S.tc[t * 4 + 2] << sh) + (l >> (2 * sh));
}
// One weighted Jacobi sweep of level k (relaxCell).
template <class St>
__device__ __forceinline__ void tcRelax(TileCycleLds &S, const Dir &dir, int C, int k, int cur, float w) {
  const int n = C << (3 * (2 - k)), off = tcOffset(C, k);
  const float *src = S.v[cur];
  float *dst = S.v[cur ^ 1];
  for (int i = (int)threadIdx.x; i < n; i += TC_LANES) {
    int t, g[3];
    tcCell(S, k, i, t, g);
    if (S.tc[t * 4 + 3] == 0) continue;
    const int a = off + i, e = S.flag[a];
    float nb[6];
    for (int d = 0; d < 6; d++) {
      const int q[3] = {g[0] + OFF[d][0], g[1] + OFF[d][1], g[2] + OFF[d][2]};
      const int j = tcIndex(S, dir, C, k, t, q);
      nb[d] = j < 0 ? 0.f : src[j];
    }
    if ((e & STORED_SOLID) || !(e & UNKNOWN)) {
      dst[a] = 0.f;
      continue;
    }
    const float cand = (openSum(nb, e & 63, w) - S.rhs[a]) / fmaxv(diagonal(w, e & 63), 1e-12f);
    dst[a] = stored<St>(mixv(src[a], cand, 0.8f));
  }
}
// The restriction from level 1 (current iterate `cur`) into the bottom, whose first iterate
// lands in v[0] (restrictCell: the eight children are the tile slot's 2^3 cells).
template <class St>
__device__ __forceinline__ void tcRestrict(TileCycleLds &S, const Dir &dir, int C, int cur, float w) {
  const float *fv = S.v[cur];
  const int stride[3] = {1, 2, 4};
  for (int t = (int)threadIdx.x; t < C; t += TC_LANES) {
    if (S.tc[t * 4 + 3] == 0) continue;
    const int id[3] = {S.tc[t * 4], S.tc[t * 4 + 1], S.tc[t * 4 + 2]};
    const int p0[3] = {id[0] * 2, id[1] * 2, id[2] * 2};
    const int a0 = t * 8;
    float side[6][4];
    for (int f = 0; f < 6; f++) {
      const int ax = f >> 1, step = (f & 1) ? -1 : 2;
      const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;
      int q[3] = {p0[0], p0[1], p0[2]};
      q[ax] += step;
      const int qa = tcIndex(S, dir, C, 1, t, q);  // a side of 2: always across the slot
      for (int j = 0; j < 4; j++) side[f][j] = qa < 0 ? 0.f : fv[qa + (j & 1) * stride[u] + (j >> 1) * stride[v]];
    }
    float residual = 0.f;
    bool every = true;
    for (int c = 0; c < 8; c++) {
      const int x = c & 1, y = (c >> 1) & 1, z = c >> 2;
      const int a = a0 + c;
      const int e = S.flag[a];
      const bool unknown = (e & UNKNOWN) != 0;
      every = every && unknown;
      if (!unknown || (e & STORED_SOLID)) continue;
      const float nb[6] = {x ? side[0][y + 2 * z] : fv[a + 1], x ? fv[a - 1] : side[1][y + 2 * z],
                           y ? side[2][x + 2 * z] : fv[a + 2], y ? fv[a - 2] : side[3][x + 2 * z],
                           z ? side[4][x + 2 * y] : fv[a + 4], z ? fv[a - 4] : side[5][x + 2 * y]};
      const int flags = e & 63;
      residual += S.rhs[a] - (openSum(nb, flags, w) - diagonal(w, flags) * fv[a]);
    }
    const float rhs = residual * 0.5f;
    const int out = tcOffset(C, 2) + t;
    S.rhs[out] = stored<St>(rhs);
    float first = 0.f;
    if (every) {
      const int e = S.flag[out];
      if (!(e & STORED_SOLID)) first = mixv(0.f, (0.f - rhs) / fmaxv(diagonal(w, e & 63), 1e-12f), 0.8f);
    }
    S.v[0][out] = stored<St>(first);
  }
}
// coarseValue at the bottom's cell cc (current iterate `cur`), from the home tile t.
__device__ __forceinline__ float tcCoarseValue(const TileCycleLds &S, const Dir &dir, const Params &P, int C, int cur,
                                               int t, const int cc0[3]) {
  const int groundY = P.ground ? 0 : (int)0x80000000;
  const int cc[3] = {cc0[0], cc0[1] > groundY ? cc0[1] : groundY, cc0[2]};
  const int a0 = tcIndex(S, dir, C, 2, t, cc);
  if (!(a0 >= 0 && (S.topo[a0] & STORED_SOLID))) return a0 < 0 ? 0.f : S.v[cur][a0];
  float sum = 0.f, count = 0.f;
  for (int axis = 0; axis < 3; axis++)
    for (int sd = -1; sd <= 1; sd += 2) {
      int q[3] = {cc[0], cc[1], cc[2]};
      q[axis] += sd;
      q[1] = q[1] > groundY ? q[1] : groundY;
      const int a = tcIndex(S, dir, C, 2, t, q);
      if (!(a >= 0 && (S.topo[a] & STORED_SOLID))) {
        sum += a < 0 ? 0.f : S.v[cur][a];
        count += 1.0f;
      }
    }
  return sum / fmaxv(count, 1.0f);
}
// The prolongation from the bottom (current iterate `ccur`) into level 1 (prolongCell).
template <class St>
__device__ __forceinline__ void tcProlong(TileCycleLds &S, const Dir &dir, const Params &P, int C, int cur, int ccur) {
  const int n = C * 8;
  const float *src = S.v[cur];
  float *dst = S.v[cur ^ 1];
  for (int i = (int)threadIdx.x; i < n; i += TC_LANES) {
    int t, p[3];
    tcCell(S, 1, i, t, p);
    if (S.tc[t * 4 + 3] == 0) continue;
    const int e = S.flag[i];
    if ((e & STORED_SOLID) || !(e & UNKNOWN)) {
      dst[i] = 0.f;
      continue;
    }
    float f[3];
    int b[3];
    for (int x = 0; x < 3; x++) {
      const float q = ((float)p[x] + 0.5f) * 0.5f - 0.5f;
      const float fl = floorf(q);
      b[x] = (int)fl;
      f[x] = q - fl;
    }
    const bool plain = P.colliders == 0 && (!P.ground || b[1] >= 0);
    float v[8];
    for (int c = 0; c < 8; c++) {
      const int cc[3] = {b[0] + (c & 1), b[1] + ((c >> 1) & 1), b[2] + (c >> 2)};
      if (plain) {
        const int j = tcIndex(S, dir, C, 2, t, cc);
        v[c] = j < 0 ? 0.f : S.v[ccur][j];
      } else {
        v[c] = tcCoarseValue(S, dir, P, C, ccur, t, cc);
      }
    }
    const float corr = mixv(mixv(mixv(v[0], v[1], f[0]), mixv(v[2], v[3], f[0]), f[1]),
                            mixv(mixv(v[4], v[5], f[0]), mixv(v[6], v[7], f[0]), f[1]), f[2]);
    dst[i] = stored<St>(src[i] + corr);
  }
}
// code: the flips (bits 0-15) and, in bit 16, the bottom iterations in FP64 (mg_bottom64)
// rather than FP32 (mg_bottom).
template <class St>
__device__ __forceinline__ void mg_coarse_cycle_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const int C = m.dir.capacity;
  if (vb != 0 || !m.tilesFused || C > TC_MAX) return;
  __shared__ TileCycleLds S;
  const bool bottom64 = ((code >> 16) & 1) != 0;
  const int L1 = m.L.brickLevels + 1;
  const auto lv1 = levelView(m, L1, code & 0xffff);
  const auto lvb = levelView(m, L1 + 1, code & 0xffff);
  const float w = m.weight;
  const int tid = (int)threadIdx.x;
  for (int i = tid; i < C * 27; i += TC_LANES) {
    const int t = i / 27;
    S.tnb[i] = m.dir.tileRec[(size_t)t * TILE_RECORD_WORDS + 4 + (i - t * 27)];
  }
  for (int i = tid; i < C * 4; i += TC_LANES) S.tc[i] = m.dir.tileRec[(size_t)(i >> 2) * TILE_RECORD_WORDS + (i & 3)];
  for (int i = tid; i < C * 8; i += TC_LANES) {
    S.v[0][i] = St::ld(lv1.p[i]);
    S.rhs[i] = St::ld(lv1.rhs[i]);
    S.flag[i] = lv1.eflag[i];
    S.topo[i] = lv1.topo[i];
  }
  for (int i = tid; i < C; i += TC_LANES) {
    S.flag[C * 8 + i] = lvb.eflag[i];
    S.topo[C * 8 + i] = lvb.topo[i];
  }
  __syncthreads();
  int cur = 0;
  for (int r = 0; r < 2; r++) {
    tcRelax<St>(S, m.dir, C, 1, cur, w);
    cur ^= 1;
    __syncthreads();
  }
  tcRestrict<St>(S, m.dir, C, cur, w);
  __syncthreads();
  // The bottom iterations (relaxCoarse), one tile slot a lane, as mg_bottom or mg_bottom64.
  {
    const int off = tcOffset(C, 2), t = tid;
    const bool in = t < C;
    const bool valid = in && S.tc[t * 4 + 3] != 0;
    const int e = valid ? S.flag[off + t] : 0;
    const bool live = valid && (e & UNKNOWN) && !(e & STORED_SOLID);
    const int flags = e & 63;
    int nbs[6];
    for (int sd = 0; sd < 6; sd++)
      nbs[sd] = live && ((flags >> sd) & 1) ? S.tnb[t * 27 + (OFF[sd][2] + 1) * 9 + (OFF[sd][1] + 1) * 3 + (OFF[sd][0] + 1)]
                                            : -1;
    const float first = valid && (e & UNKNOWN) ? S.v[0][off + t] : 0.f;
    const float rhs = valid ? S.rhs[off + t] : 0.f;
    float result;
    if (!bottom64) {
      float *bv[2] = {S.v[0] + off, S.v[1] + off};
      __syncthreads();  // every lane has read the restricted iterate
      if (in) bv[0][t] = first;
      __syncthreads();
      float mine = first;
      for (int it = 0; it < BOTTOM_ITERATIONS; it++) {
        const float *cv = bv[it & 1];
        float out = 0.f;
        if (live) {
          float nb[6];
          for (int sd = 0; sd < 6; sd++) nb[sd] = nbs[sd] >= 0 ? cv[nbs[sd]] : 0.f;
          const float sum = w * (nb[0] + nb[1]) + w * (nb[2] + nb[3]) + w * (nb[4] + nb[5]);
          const float cand = (sum - rhs) / fmaxv(diagonal(w, flags), 1e-12f);
          out = mixv(mine, cand, 0.8f);
        }
        mine = out;
        if (in) bv[(it & 1) ^ 1][t] = out;
        __syncthreads();
      }
      result = mine;
    } else {
      const double wd = (double)w;
      double mine = (double)first;
      if (in) S.b64[0][t] = mine;
      __syncthreads();
      for (int it = 0; it < BOTTOM_ITERATIONS; it++) {
        const double *cv = S.b64[it & 1];
        double out = 0.0;
        if (live) {
          double nb[6];
          for (int sd = 0; sd < 6; sd++) nb[sd] = nbs[sd] >= 0 ? cv[nbs[sd]] : 0.0;
          const double sum = wd * (nb[0] + nb[1]) + wd * (nb[2] + nb[3]) + wd * (nb[4] + nb[5]);
          const double diag = wd * (double)openCount(flags, 0) + wd * (double)openCount(flags, 2) +
                              wd * (double)openCount(flags, 4);
          const double cand = (sum - (double)rhs) / (diag < 1e-12 ? 1e-12 : diag);
          out = mix64(mine, cand, 0.8);
        }
        mine = out;
        if (in) S.b64[(it & 1) ^ 1][t] = out;
        __syncthreads();
      }
      result = (float)mine;
    }
    __syncthreads();
    if (valid) S.v[0][off + t] = stored<St>(result);
    __syncthreads();
  }
  tcProlong<St>(S, m.dir, m.P, C, cur, 0);
  cur ^= 1;
  __syncthreads();
  for (int r = 0; r < 3; r++) {
    tcRelax<St>(S, m.dir, C, 1, cur, w);
    cur ^= 1;
    __syncthreads();
  }
  for (int i = tid; i < C * 8; i += TC_LANES)
    if (S.tc[(i >> 3) * 4 + 3] != 0) lv1.p[i] = St::st(S.v[cur][i]);
}
FP_MG_PAIR(mg_coarse_cycle, 1024)

// ---- defect correction -----------------------------------------------------------------
// The loop's start, one workgroup per pool slot: x = 0 on the slot's level-0 cells; the
// first workgroup clears the finished flag and the cycle count.
template <class St> __device__ __forceinline__ void dc_start_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  (void)code;
  const int sd = m.lv[0].side, n = sd * sd * sd;
  if (vb < m.dcSlots)
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
      m.dcX[(size_t)vb * n + l] = 0.f;
      m.dcXlo[(size_t)vb * n + l] = 0.f;
    }
  if (vb == 0 && threadIdx.x == 0) {
    m.dcNorms[4] = 1.0;
    m.dcNorms[5] = 0.0;
    m.dcNorms[6] = 0.0;
  }
}
template <class St> __global__ void __launch_bounds__(256) k_dc_start(MGArgsT<typename St::T> a, int code) {
  dc_start_body<St>(a, (int)blockIdx.x, code);
}
template <class St>
__global__ void __launch_bounds__(256)
    k_dc_start_batch(const MGArgsT<typename St::T> *args, const int *starts, int n, int code) {
  FP_BATCH_LOOP(dc_start, dc_start_body<St>(args[sim], v - starts[sim], code))
}

// The level-0 operator of relaxCell as a residual: r = b - (sum - diag x) on unknown cells,
// zero elsewhere, each row's dot product accumulated in FP64 from x = hi + lo and rounded
// to FP32; per-brick FP64 sums of r^2 and b^2 and the largest |r|.
template <class St>
__device__ __forceinline__ void dc_residual_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  (void)code;
  const auto &lv = m.lv[0];
  const int sd = lv.side, n = sd * sd * sd;
  MG_BRICK_GROUP(n)
  __shared__ double red[3][256];
  __shared__ int recs_[MG_GROUP][HOME_WORDS];
  loadGroupHomes(m.dir, m.active, count_, gr, recs_, m.L.side);
  double rr = 0.0, bb = 0.0;
  float big = 0.f;
  if (mine) {
    const int id = m.active[bi];
    MGCtx c;
    c.dir = m.dir;
    c.L = m.L;
    c.side = m.L.side;
    c.hasBrick = true;
    brickTile(c, id);
    c.home = homeOf(m.dir, id, recs_[gr.j]);
    const int slot = entrySlot(recs_[gr.j][13], c.side);
    const float *x = m.dcX;
    for (int l = gr.first; l < n; l += gr.step) {
      size_t a = (size_t)slot * n + l;
      int e = lv.eflag[a];
      float v = 0.f;
      if (!(e & STORED_SOLID) && (e & UNKNOWN)) {
        int p[3];
        levelCell(c.home.bx, c.home.by, c.home.bz, sd, l, p);
        int flags = e & 63;
        const double w = (double)m.weight;
        double sum = 0.0;
        const int msd = sd - 1;
        for (int k = 0; k < 6; k++) {
          if (!((flags >> k) & 1)) continue;
          const int lx = (p[0] & msd) + OFF[k][0], ly = (p[1] & msd) + OFF[k][1], lz = (p[2] & msd) + OFF[k][2];
          double xv;
          if ((unsigned)lx < (unsigned)sd && (unsigned)ly < (unsigned)sd && (unsigned)lz < (unsigned)sd) {
            const size_t b = a + OFF[k][0] + (OFF[k][1] + OFF[k][2] * sd) * sd;
            xv = (double)x[b] + (double)m.dcXlo[b];
          } else {
            int q[3] = {p[0] + OFF[k][0], p[1] + OFF[k][1], p[2] + OFF[k][2]};
            xv = (double)readAt<StoreF32>(c, x, q, 0) + (double)readAt<StoreF32>(c, m.dcXlo, q, 0);
          }
          sum += w * xv;
        }
        float bv = m.dcB[a];
        v = (float)((double)bv - (sum - (double)diagonal(m.weight, flags) * ((double)x[a] + (double)m.dcXlo[a])));
        rr += (double)v * (double)v;
        bb += (double)bv * (double)bv;
        big = fmaxv(big, fabsv(v));
      }
      m.dcR[a] = v;
    }
  }
  // The group's sums: one partial per group, in a fixed order.
  red[0][threadIdx.x] = rr;
  red[1][threadIdx.x] = bb;
  red[2][threadIdx.x] = (double)big;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] = red[2][threadIdx.x] < red[2][threadIdx.x + s] ? red[2][threadIdx.x + s] : red[2][threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    m.dcPartials[3 * gr.g] = red[0][0];
    m.dcPartials[3 * gr.g + 1] = red[1][0];
    m.dcPartials[3 * gr.g + 2] = red[2][0];
  }
}
FP_MG_PAIR(dc_residual, 256)
// The totals of the per-brick sums in a fixed order: (sum r^2, sum b^2, max |r|, scale), the
// scale being the largest power of two not above max |r| (1 for zero or subnormal
// residuals). Scaled by it, the correction equation's values lie in (-2, 2), inside every
// storage format's range.
template <class St> __device__ __forceinline__ void dc_norms_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  (void)code;
  if (vb != 0) return;
  __shared__ double red[3][256];
  const int n0 = m.lv[0].side * m.lv[0].side * m.lv[0].side;
  const int k0 = groupSize(n0, 256);  // the residual's groups (256 lanes)
  const int blocks = (m.counts[0] + k0 - 1) / k0;
  double rr = 0.0, bb = 0.0, big = 0.0;
  for (int i = threadIdx.x; i < blocks; i += 256) {
    rr += m.dcPartials[3 * i];
    bb += m.dcPartials[3 * i + 1];
    big = big < m.dcPartials[3 * i + 2] ? m.dcPartials[3 * i + 2] : big;
  }
  red[0][threadIdx.x] = rr;
  red[1][threadIdx.x] = bb;
  red[2][threadIdx.x] = big;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
      red[2][threadIdx.x] = red[2][threadIdx.x] < red[2][threadIdx.x + s] ? red[2][threadIdx.x + s] : red[2][threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    m.dcNorms[0] = red[0][0];
    m.dcNorms[1] = red[1][0];
    m.dcNorms[2] = red[2][0];
    uint32_t bits = __float_as_uint((float)red[2][0]) & 0x7f800000u;
    m.dcNorms[3] = bits ? (double)__uint_as_float(bits) : 1.0;
    // The stopping test after `code` cycles: |r| / |b| at most the tolerance, or the
    // cycle budget spent.
    const double rel = red[1][0] > 0.0 ? sqrt(red[0][0] / red[1][0]) : 0.0;
    m.dcNorms[4] = rel;
    if (rel <= m.dcTol || code >= m.dcMaxCycles) {
      m.dcNorms[6] = (double)code;
      m.dcNorms[5] = 1.0;
    }
  }
}
FP_MG_PAIR(dc_norms, 256)
// Level 0's right-hand side: r / scale, stored.
template <class St> __device__ __forceinline__ void dc_scale_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  (void)code;
  const auto &lv = m.lv[0];
  const int n = lv.side * lv.side * lv.side;
  MG_BRICK_GROUP(n)
  if (!mine) return;
  const int id = m.active[bi];
  const int slot = entrySlot(m.dir.entries[id], m.L.side);
  const float inv = 1.0f / (float)m.dcNorms[3];
  for (int l = gr.first; l < n; l += gr.step) {
    size_t a = (size_t)slot * n + l;
    lv.rhs[a] = St::st(m.dcR[a] * inv);
  }
}
FP_MG_PAIR(dc_scale, 256)
// x += scale * e, level 0's solution of the scaled correction equation, on unknown cells.
// x is the pair hi + lo: the sum hi + v is split exactly into its rounded value and error
// (Knuth's two-sum), the error joins lo, and the pair is renormalized so that hi is the
// rounded value of hi + lo. A single FP32 x could not hold the solution closer than about
// the unit roundoff times the operator's condition number, and the residual would level
// off there.
__device__ __forceinline__ void addToPair(float &hi, float &lo, float v) {
  const float s = hi + v;
  const float bp = s - hi;
  const float err = (hi - (s - bp)) + (v - bp);
  const float l = lo + err;
  const float h = s + l;
  lo = l - (h - s);
  hi = h;
}
template <class St> __device__ __forceinline__ void dc_update_body(const MGArgsT<typename St::T> &m, const int vb, int code) {
  const auto lv = levelView(m, 0, code >> 8);
  const int n = lv.side * lv.side * lv.side;
  MG_BRICK_GROUP(n)
  if (!mine) return;
  const int id = m.active[bi];
  const int slot = entrySlot(m.dir.entries[id], m.L.side);
  const float scale = (float)m.dcNorms[3];
  for (int l = gr.first; l < n; l += gr.step) {
    size_t a = (size_t)slot * n + l;
    int e = lv.eflag[a];
    float hi = 0.f, lo = 0.f;
    if (!(e & STORED_SOLID) && (e & UNKNOWN)) {
      hi = m.dcX[a];
      lo = m.dcXlo[a];
      addToPair(hi, lo, scale * St::ld(lv.p[a]));
    }
    m.dcX[a] = hi;
    m.dcXlo[a] = lo;
  }
}
FP_MG_PAIR(dc_update, 256)

// The batched solve's virtual blocks, per level: for member i, ceil(units / k) groups
// rounded up to a multiple of 8 (units: active bricks at a brick level, tile slots at a
// tile level; k: groupSize at 256 lanes for brick levels, 64 for tile levels), or none
// once its defect correction has finished; out[MAX_LEVELS] holds one block per
// unfinished member (the per-simulation kernels). Prefix sums, `n` members, each array
// n + 1 entries; one workgroup of 1024 lanes.
constexpr int PREFIX_LANES = 1024;
template <class St>
__device__ __forceinline__ int levelGroups(const MGArgsT<typename St::T> &a, int L, int cells, bool tile) {
  if (!a.counts || dcFinished(a)) return 0;
  if (L == MAX_LEVELS) return 1;
  // The levels below the first tile level of such a member run in k_mg_coarse_cycle_batch.
  if (tile && a.tilesFused && L > a.L.brickLevels) return 0;
  const int units = tile ? a.dir.capacity : a.counts[0];
  const int k = groupSize(cells, tile ? 64 : 256);
  return ((units + k - 1) / k + 7) / 8 * 8;
}
template <class St>
__global__ void __launch_bounds__(PREFIX_LANES) k_level_prefix_batch(const MGArgsT<typename St::T> *args, int n, int *out) {
  __shared__ int waveTotals[PREFIX_LANES / 64];
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const int per = (n + PREFIX_LANES - 1) / PREFIX_LANES;
  const int lo = (int)threadIdx.x * per, hi = lo + per < n ? lo + per : n;
  const int bottom = n > 0 && args[0].counts ? args[0].bottom : 0;
  for (int L = 0; L <= MAX_LEVELS; L++) {
    if (L < MAX_LEVELS && L > bottom) continue;
    // Level shapes are the same for every member.
    const int sd = L < MAX_LEVELS ? args[0].lv[L].side : 1, cells = sd * sd * sd;
    const bool tile = L < MAX_LEVELS && args[0].lv[L].tile != 0;
    int mine = 0;
    for (int i = lo; i < hi; i++) mine += levelGroups<St>(args[i], L, cells, tile);
    // Inclusive scan within the wavefront, then the wavefronts' totals.
    int v = mine;
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(v, d, 64);
      if (lane >= d) v += t;
    }
    if (lane == 63) waveTotals[wave] = v;
    __syncthreads();
    int run = v - mine;
    for (int w = 0; w < wave; w++) run += waveTotals[w];
    int *o = out + (size_t)L * (n + 1);
    for (int i = lo; i < hi; i++) {
      o[i] = run;
      run += levelGroups<St>(args[i], L, cells, tile);
    }
    if (threadIdx.x == PREFIX_LANES - 1) o[n] = run;
    __syncthreads();
  }
}
template __global__ void k_level_prefix_batch<StoreF32>(const MGArgsT<float> *, int, int *);
template __global__ void k_level_prefix_batch<StoreBF16>(const MGArgsT<uint16_t> *, int, int *);
template __global__ void k_level_prefix_batch<StoreF16>(const MGArgsT<_Float16> *, int, int *);

#define FP_MG_INSTANCE(NAME, St)                                         \
  template __global__ void k_##NAME<St>(MGArgsT<St::T>, int);           \
  template __global__ void k_##NAME##_batch<St>(const MGArgsT<St::T> *, const int *, int, int);
#define FP_MG_INSTANTIATE(St)              \
  FP_MG_INSTANCE(mg_init, St)              \
  FP_MG_INSTANCE(mg_relax_bricks, St)      \
  FP_MG_INSTANCE(mg_relax_tiles, St)       \
  FP_MG_INSTANCE(mg_restrict_bricks, St)   \
  FP_MG_INSTANCE(mg_restrict_tiles, St)    \
  FP_MG_INSTANCE(mg_prolong_bricks, St)    \
  FP_MG_INSTANCE(mg_prolong_tiles, St)     \
  FP_MG_INSTANCE(mg_bottom, St)            \
  FP_MG_INSTANCE(mg_bottom64, St)          \
  FP_MG_INSTANCE(mg_tiles_down, St)        \
  FP_MG_INSTANCE(mg_tiles_up, St)          \
  FP_MG_INSTANCE(mg_bricks_down, St)       \
  FP_MG_INSTANCE(mg_bricks_up, St)         \
  FP_MG_INSTANCE(mg_coarse_cycle, St)      \
  FP_MG_INSTANCE(dc_start, St)             \
  FP_MG_INSTANCE(dc_residual, St)          \
  FP_MG_INSTANCE(dc_norms, St)             \
  FP_MG_INSTANCE(dc_scale, St)             \
  FP_MG_INSTANCE(dc_update, St)
FP_MG_INSTANTIATE(StoreF32)
FP_MG_INSTANTIATE(StoreBF16)
FP_MG_INSTANTIATE(StoreF16)

}  // namespace fp
