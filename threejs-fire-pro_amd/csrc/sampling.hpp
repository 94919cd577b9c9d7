// Cell access and trilinear sampling over the sparse brick pools, for kernels whose work
// item lies in one active brick (the home brick, whose record is staged in LDS).
#pragma once
#include "device.hpp"

namespace fp {

// Device pools, slot-major. Field cells are float4 (smoke, heat, lifetime, fuel); velocity
// is three planes per slot (the MAC faces +x, +y, +z of each cell).
struct Pools {
  float4 *F, *Fp, *Fn;
  // The scalar backtrace of each field cell, left by advection for the corrector: three
  // planes of float4 per slot, (forward departure point, lifetime min), (smoke, heat, fuel
  // min, lifetime max), (smoke, heat, fuel max, content support bits).
  float4 *Don;
  // The velocity backtrace of each velocity cell, left by advection for the corrector: four
  // planes of float4 per slot, (face velocity, donor min) of components x, y, z, then the
  // three donor maxima.
  float4 *Vtr;
  // Coarse smoke density, predictor and next (smokeDivisor 2 or 4), (B / sd)^3 per slot.
  float *S, *Sp, *Sn;
  float *E;
  float *V, *Vadv, *Vcor, *Vf;
  float4 *Curl;
  float *D;
  uint32_t *solids;
  const float4 *noise;
};

struct Ctx {
  Home home;
  Dir dir;
  Layout L;
  int side;
};

__device__ __forceinline__ int slotAtBrick(const Ctx &c, int bx, int by, int bz) {
  int dx = bx - c.home.bx, dy = by - c.home.by, dz = bz - c.home.bz;
  if ((unsigned)(dx + 1) <= 2u && (unsigned)(dy + 1) <= 2u && (unsigned)(dz + 1) <= 2u)
    return c.home.slot[(dz + 1) << 4 | (dy + 1) << 2 | (dx + 1)];
  return entrySlot(entryLookup(c.dir, bx, by, bz), c.side);
}

// Brick edges are powers of two: a cell's index inside its brick and a slot's first cell
// are shifts (sh the edge's log2), not multiplies, which are quarter rate on the VALU.
__device__ __forceinline__ int cellIndex(int sh, int x, int y, int z) {
  const int m = (1 << sh) - 1;
  return (((z & m) << sh | (y & m)) << sh) | (x & m);
}
__device__ __forceinline__ size_t slotBase(int slot, int sh) { return (size_t)(uint32_t)slot << (3 * sh); }

__device__ __forceinline__ float4 fieldCell(const Ctx &c, const float4 *buf, int x, int y, int z) {
  int s = slotAtBrick(c, x >> c.L.sF, y >> c.L.sF, z >> c.L.sF);
  if (s < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
  return buf[slotBase(s, c.L.sF) + cellIndex(c.L.sF, x, y, z)];
}
__device__ __forceinline__ float fieldScalar(const Ctx &c, const float *buf, int x, int y, int z) {
  int s = slotAtBrick(c, x >> c.L.sF, y >> c.L.sF, z >> c.L.sF);
  if (s < 0) return 0.f;
  return buf[slotBase(s, c.L.sF) + cellIndex(c.L.sF, x, y, z)];
}
__device__ __forceinline__ float velCell(const Ctx &c, const float *buf, int comp, int x, int y, int z) {
  int s = slotAtBrick(c, x >> c.L.sV, y >> c.L.sV, z >> c.L.sV);
  if (s < 0) return 0.f;
  return buf[slotBase(s * 3 + comp, c.L.sV) + cellIndex(c.L.sV, x, y, z)];
}
__device__ __forceinline__ float velScalar(const Ctx &c, const float *buf, int x, int y, int z) {
  int s = slotAtBrick(c, x >> c.L.sV, y >> c.L.sV, z >> c.L.sV);
  if (s < 0) return 0.f;
  return buf[slotBase(s, c.L.sV) + cellIndex(c.L.sV, x, y, z)];
}
__device__ __forceinline__ float4 curlCell(const Ctx &c, const float4 *buf, int x, int y, int z) {
  int s = slotAtBrick(c, x >> c.L.sV, y >> c.L.sV, z >> c.L.sV);
  if (s < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
  return buf[slotBase(s, c.L.sV) + cellIndex(c.L.sV, x, y, z)];
}

// ---- accesses within the home brick's neighborhood ---------------------------------------
// Cells at most two cells from the work item's brick (the force, divergence and projection
// kernels' reads, every kernel's solid test of its own cell) lie inside the brick's 3^3
// neighborhood: their slots come from the LDS table alone, and every lane loads (from slot 0
// where the brick has none, then zero), so lanes take the same instructions whatever brick
// a cell lies in.
__device__ __forceinline__ int nearSlot(const Ctx &c, int bx, int by, int bz) {
  return c.home.slot[((bz - c.home.bz + 1) << 4) | ((by - c.home.by + 1) << 2) | (bx - c.home.bx + 1)];
}
__device__ __forceinline__ float velCellNear(const Ctx &c, const float *buf, int comp, int x, int y, int z) {
  const int s = nearSlot(c, x >> c.L.sV, y >> c.L.sV, z >> c.L.sV);
  const float v = buf[slotBase((s < 0 ? 0 : s) * 3 + comp, c.L.sV) + cellIndex(c.L.sV, x, y, z)];
  return s < 0 ? 0.f : v;
}
__device__ __forceinline__ float velScalarNear(const Ctx &c, const float *buf, int x, int y, int z) {
  const int s = nearSlot(c, x >> c.L.sV, y >> c.L.sV, z >> c.L.sV);
  const float v = buf[slotBase(s < 0 ? 0 : s, c.L.sV) + cellIndex(c.L.sV, x, y, z)];
  return s < 0 ? 0.f : v;
}
__device__ __forceinline__ float4 curlCellNear(const Ctx &c, const float4 *buf, int x, int y, int z) {
  const int s = nearSlot(c, x >> c.L.sV, y >> c.L.sV, z >> c.L.sV);
  const float4 v = buf[slotBase(s < 0 ? 0 : s, c.L.sV) + cellIndex(c.L.sV, x, y, z)];
  return s < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : v;
}
__device__ __forceinline__ float fieldScalarNear(const Ctx &c, const float *buf, int x, int y, int z) {
  const int s = nearSlot(c, x >> c.L.sF, y >> c.L.sF, z >> c.L.sF);
  const float v = buf[slotBase(s < 0 ? 0 : s, c.L.sF) + cellIndex(c.L.sF, x, y, z)];
  return s < 0 ? 0.f : v;
}
__device__ __forceinline__ float4 fieldCellNear(const Ctx &c, const float4 *buf, int x, int y, int z) {
  const int s = nearSlot(c, x >> c.L.sF, y >> c.L.sF, z >> c.L.sF);
  const float4 v = buf[slotBase(s < 0 ? 0 : s, c.L.sF) + cellIndex(c.L.sF, x, y, z)];
  return s < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : v;
}

// The 2x2x2 cells from (x, y, z) on a grid of bricks 2^sh cells a side, corner k = (z * 2 +
// y) * 2 + x: each corner's pool slot (-1 without one) and its index inside the slot. The
// bricks are resolved once per axis; when all of them lie in the home brick's 3^3
// neighborhood (nearly always) the slots come from the LDS record without any hashing.
__device__ __forceinline__ void cornerCells(const Ctx &c, int sh, int x, int y, int z, int slot[8], int local[8]) {
  const int n = 1 << sh, m = n - 1;
  const int p[3] = {x, y, z};
  const int home[3] = {c.home.bx, c.home.by, c.home.bz};
  const int lsh[3] = {0, sh, 2 * sh};
  int bk[3][2], rel[3][2], loc[3][2];
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int q = p[a] + e;
      bk[a][e] = q >> sh;
      const int d = bk[a][e] - home[a] + 1;
      in = in && (unsigned)d <= 2u;
      rel[a][e] = d << (2 * a);
      loc[a][e] = (q & m) << lsh[a];
    }
#pragma unroll
  for (int k = 0; k < 8; k++) local[k] = loc[2][k >> 2] + loc[1][(k >> 1) & 1] + loc[0][k & 1];
  if (in) {
#pragma unroll
    for (int k = 0; k < 8; k++) slot[k] = c.home.slot[rel[2][k >> 2] + rel[1][(k >> 1) & 1] + rel[0][k & 1]];
  } else {
    for (int k = 0; k < 8; k++) slot[k] = slotAtBrick(c, bk[0][k & 1], bk[1][(k >> 1) & 1], bk[2][k >> 2]);
  }
}
// The eight field cells from (x, y, z) (zero for cells without a slot).
__device__ __forceinline__ void fieldCorners(const Ctx &c, const float4 *buf, int x, int y, int z, float4 v[8]) {
  int slot[8], local[8];
  cornerCells(c, c.L.sF, x, y, z, slot, local);
#pragma unroll
  for (int k = 0; k < 8; k++)
    v[k] = slot[k] < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : buf[slotBase(slot[k], c.L.sF) + local[k]];
}
// The eight velocity cells of plane `comp` from (x, y, z).
__device__ __forceinline__ void velCorners(const Ctx &c, const float *buf, int comp, int x, int y, int z, float v[8]) {
  int slot[8], local[8];
  cornerCells(c, c.L.sV, x, y, z, slot, local);
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = slot[k] < 0 ? 0.f : buf[slotBase(slot[k] * 3 + comp, c.L.sV) + local[k]];
}

// The trilinear sum of eight cells, corner k = (z * 2 + y) * 2 + x weighted by
// (wx * wy) * wz, accumulated in corner order.
__device__ __forceinline__ float trilinear(const float cv[8], float fx, float fy, float fz) {
  float out = 0.f;
#pragma unroll
  for (int z = 0; z < 2; z++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int x = 0; x < 2; x++) {
        float wx = x ? fx : 1.0f - fx, wy = y ? fy : 1.0f - fy, wz = z ? fz : 1.0f - fz;
        out += cv[(z * 2 + y) * 2 + x] * (wx * wy * wz);
      }
  return out;
}
// The eight smoke cells from (x, y, z) of a coarse smoke pool.
__device__ __forceinline__ void smokeCorners(const Ctx &c, const float *buf, int x, int y, int z, float v[8]) {
  int slot[8], local[8];
  cornerCells(c, c.L.sS, x, y, z, slot, local);
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = slot[k] < 0 ? 0.f : buf[slotBase(slot[k], c.L.sS) + local[k]];
}
// A trilinear sample of a smoke pool at (px, py, pz) in smoke cells, with the range of the
// eight cells it weighs: the minimum from 1e30 and the maximum from 0, as smokeDonorBounds
// (smoke.wgsl:59-73) scans them.
__device__ __forceinline__ float sampleSmoke(const Ctx &c, const float *buf, float px, float py, float pz,
                                             float &lo, float &hi) {
  float qx = px - 0.5f, qy = py - 0.5f, qz = pz - 0.5f;
  float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
  float fx = qx - flx, fy = qy - fly, fz = qz - flz;
  float cv[8];
  smokeCorners(c, buf, (int)flx, (int)fly, (int)flz, cv);
  lo = 1e30f;
  hi = 0.f;
  range8(cv, lo, hi);
  return trilinear(cv, fx, fy, fz);
}

// solidCell / solidAt (solids.wgsl:15-36).
__device__ __forceinline__ bool solidCell(const Ctx &c, const Params &P, const uint32_t *solids, int x, int y,
                                          int z) {
  if (P.ground && y < 0) return true;
  if (P.colliders == 0 || !inSolidBox(P, x, y, z)) return false;
  int s = slotAtBrick(c, x >> c.L.sF, y >> c.L.sF, z >> c.L.sF);
  if (s < 0) return false;
  int m = c.L.B - 1, sh = c.L.sF;
  uint32_t bit = (uint32_t)(((((z & m) << sh) | (y & m)) << sh) | (x & m));
  return ((solids[(size_t)s * (c.L.nF >> 5) + (bit >> 5)] >> (bit & 31u)) & 1u) != 0u;
}
__device__ __forceinline__ bool solidAt(const Ctx &c, const Params &P, const uint32_t *solids, float wx,
                                        float wy, float wz) {
  int x = (int)floorf(wx * P.invHf + 0.25f), y = (int)floorf(wy * P.invHf + 0.25f),
      z = (int)floorf(wz * P.invHf + 0.25f);
  return solidCell(c, P, solids, x, y, z);
}

// solidAt for a cell within the home brick's neighborhood (nearSlot).
__device__ __forceinline__ bool solidAtNear(const Ctx &c, const Params &P, const uint32_t *solids, float wx, float wy,
                                            float wz) {
  const int x = (int)floorf(wx * P.invHf + 0.25f), y = (int)floorf(wy * P.invHf + 0.25f),
            z = (int)floorf(wz * P.invHf + 0.25f);
  if (P.ground && y < 0) return true;
  if (P.colliders == 0 || !inSolidBox(P, x, y, z)) return false;
  const int s = nearSlot(c, x >> c.L.sF, y >> c.L.sF, z >> c.L.sF);
  if (s < 0) return false;
  const int m = c.L.B - 1, sh = c.L.sF;
  const uint32_t bit = (uint32_t)(((((z & m) << sh) | (y & m)) << sh) | (x & m));
  return ((solids[(size_t)s * (c.L.nF >> 5) + (bit >> 5)] >> (bit & 31u)) & 1u) != 0u;
}
// sampleField over cells within the home brick's neighborhood (nearSlot).
__device__ __forceinline__ float4 sampleFieldNear(const Ctx &c, const float4 *buf, float px, float py, float pz) {
  float qx = px - 0.5f, qy = py - 0.5f, qz = pz - 0.5f;
  float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
  int bx = (int)flx, by = (int)fly, bz = (int)flz;
  float fx = qx - flx, fy = qy - fly, fz = qz - flz;
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 cv[8];
#pragma unroll
  for (int k = 0; k < 8; k++) cv[k] = fieldCellNear(c, buf, bx + (k & 1), by + ((k >> 1) & 1), bz + (k >> 2));
#pragma unroll
  for (int z = 0; z < 2; z++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int x = 0; x < 2; x++) {
        float wx = x ? fx : 1.0f - fx, wy = y ? fy : 1.0f - fy, wz = z ? fz : 1.0f - fz;
        float w = wx * wy * wz;
        float4 v = cv[(z * 2 + y) * 2 + x];
        out.x += v.x * w;
        out.y += v.y * w;
        out.z += v.z * w;
        out.w += v.w * w;
      }
  return out;
}

// Trilinear sample of the field grid, position in cell units with centers at .5
// (fluid-sampling.wgsl:235-264): cell values weighted by (wx * wy) * wz.
__device__ __forceinline__ float4 sampleField(const Ctx &c, const float4 *buf, float px, float py, float pz) {
  float qx = px - 0.5f, qy = py - 0.5f, qz = pz - 0.5f;
  float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
  int bx = (int)flx, by = (int)fly, bz = (int)flz;
  float fx = qx - flx, fy = qy - fly, fz = qz - flz;
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 cv[8];
  fieldCorners(c, buf, bx, by, bz, cv);
#pragma unroll
  for (int z = 0; z < 2; z++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int x = 0; x < 2; x++) {
        float wx = x ? fx : 1.0f - fx, wy = y ? fy : 1.0f - fy, wz = z ? fz : 1.0f - fz;
        float w = wx * wy * wz;
        float4 v = cv[(z * 2 + y) * 2 + x];
        out.x += v.x * w;
        out.y += v.y * w;
        out.z += v.z * w;
        out.w += v.w * w;
      }
  return out;
}
__device__ __forceinline__ float sampleVel(const Ctx &c, const float *buf, int comp, float px, float py,
                                           float pz) {
  float qx = px - 0.5f, qy = py - 0.5f, qz = pz - 0.5f;
  float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
  int bx = (int)flx, by = (int)fly, bz = (int)flz;
  float fx = qx - flx, fy = qy - fly, fz = qz - flz;
  float cv[8];
  velCorners(c, buf, comp, bx, by, bz, cv);
  return trilinear(cv, fx, fy, fz);
}
// velocityAt (fluid-advection.wgsl:10-21).
__device__ __forceinline__ void velocityAt(const Ctx &c, const Params &P, const float *V, float wx, float wy,
                                           float wz, float out[3]) {
  const float cx = wx * P.invHv, cy = wy * P.invHv, cz = wz * P.invHv;
  out[0] = sampleVel(c, V, 0, (wx - 0.5f * P.hv) * P.invHv, cy, cz);
  out[1] = sampleVel(c, V, 1, cx, (wy - 0.5f * P.hv) * P.invHv, cz);
  out[2] = sampleVel(c, V, 2, cx, cy, (wz - 0.5f * P.hv) * P.invHv);
}
// ---- the velocity window ------------------------------------------------------------------
// The field cells of one workgroup, a z-slab of its brick, sample the velocity planes close
// to that slab: at rest a cell's trilinear footprint reaches one velocity cell past the
// slab's own, and a midpoint trace moves it by 0.5 dt |v| / hv cells. stageVelWindow copies
// the slab's velocity cells and VW_HALO more on every side, all three planes, into LDS once
// per workgroup (zero where a brick has no slot, as velCorners reads such cells); a sample
// whose eight cells lie inside reads them there, any other one reads the pools. The values,
// and so every result, are the same either way: the window saves each corner's slot
// resolution, its 64-bit address and its global load.
constexpr int VW_HALO = 2;
constexpr int VW_CELLS = 768;  // 9 KB of LDS: bricks of 8 and 16 cells at velocityDivisor 2 fit
struct VelWindow {
  const float *v;  // [3][n], x fastest
  int o[3];        // the velocity cell at window cell (0, 0, 0)
  int e[3];        // extents; {1, 1, 1} when the window is off, so that no sample lies inside
  int sy, sz, n;   // row and layer strides, cells per plane
};
__device__ __forceinline__ void velWindowOff(float *lds, VelWindow &w) {
  w.v = lds;
  w.o[0] = w.o[1] = w.o[2] = 0;
  w.e[0] = w.e[1] = w.e[2] = 1;
  w.sy = w.sz = 1;
  w.n = 0;
}
// The window of the home brick's field cells with local z in [zlo, zhi]. Every thread of the
// workgroup calls it with the same arguments (it may end in a barrier).
__device__ __forceinline__ void stageVelWindow(const Ctx &c, const float *V, int zlo, int zhi, float *lds,
                                               VelWindow &w) {
  const int b = c.L.b, sV = c.L.sV, dsh = c.L.sF - c.L.sV;
  const int ex = b + 2 * VW_HALO, ez = (zhi >> dsh) - (zlo >> dsh) + 1 + 2 * VW_HALO;
  const int n = ex * ex * ez;
  // The halo must stay inside the home brick's 3^3 neighborhood, whose slots are staged.
  if (b < VW_HALO || n > VW_CELLS) {
    velWindowOff(lds, w);
    return;
  }
  w.v = lds;
  w.o[0] = c.home.bx * b - VW_HALO;
  w.o[1] = c.home.by * b - VW_HALO;
  w.o[2] = c.home.bz * b + (zlo >> dsh) - VW_HALO;
  w.e[0] = w.e[1] = ex;
  w.e[2] = ez;
  w.sy = ex;
  w.sz = ex * ex;
  w.n = n;
  // Quotients by ex as a multiply and a shift: with inv = ceil(2^16 / ex), i * inv / 2^16
  // exceeds i / ex by less than i / 2^16, which stays below 1 / ex while i * ex < 2^16
  // (here i < VW_CELLS and ex * ex <= VW_CELLS, so i * ex < 768 * 28).
  const uint32_t inv = (65536u + (uint32_t)ex - 1u) / (uint32_t)ex;
  // Four cells a lane per pass, all their loads issued before the first LDS store, so a
  // lane waits for memory once per pass rather than once per cell.
  constexpr int U = 4;
  const int lanes = (int)blockDim.x;
  for (int first = threadIdx.x; first < n; first += U * lanes) {
    float v[U][3];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int i = first + j * lanes;
      v[j][0] = v[j][1] = v[j][2] = 0.f;
      if (i < n) {
        const int t = (int)(__umul24((uint32_t)i, inv) >> 16);
        const int u = (int)(__umul24((uint32_t)t, inv) >> 16);
        const int x = w.o[0] + (i - t * ex), y = w.o[1] + (t - u * ex), z = w.o[2] + u;
        const int s = c.home.slot[((z >> sV) - c.home.bz + 1) << 4 | ((y >> sV) - c.home.by + 1) << 2 |
                                  ((x >> sV) - c.home.bx + 1)];
        const int local = cellIndex(sV, x, y, z);
        if (s >= 0) {
#pragma unroll
          for (int k = 0; k < 3; k++) v[j][k] = V[slotBase(s * 3 + k, sV) + local];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int i = first + j * lanes;
      if (i < n) {
#pragma unroll
        for (int k = 0; k < 3; k++) lds[k * n + i] = v[j][k];
      }
    }
  }
  __syncthreads();
}
// sampleVel of the pool the window was staged from.
__device__ __forceinline__ float windowVel(const Ctx &c, const VelWindow &w, const float *V, int comp, float px,
                                           float py, float pz) {
  float qx = px - 0.5f, qy = py - 0.5f, qz = pz - 0.5f;
  float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
  int bx = (int)flx, by = (int)fly, bz = (int)flz;
  float fx = qx - flx, fy = qy - fly, fz = qz - flz;
  float cv[8];
  const int rx = bx - w.o[0], ry = by - w.o[1], rz = bz - w.o[2];
  if ((unsigned)rx < (unsigned)(w.e[0] - 1) && (unsigned)ry < (unsigned)(w.e[1] - 1) &&
      (unsigned)rz < (unsigned)(w.e[2] - 1)) {
    const float *q = w.v + __mul24(comp, w.n) + __mul24(rz, w.sz) + __mul24(ry, w.sy) + rx;
    const int sy = w.sy, sz = w.sz;
    cv[0] = q[0];
    cv[1] = q[1];
    cv[2] = q[sy];
    cv[3] = q[sy + 1];
    cv[4] = q[sz];
    cv[5] = q[sz + 1];
    cv[6] = q[sz + sy];
    cv[7] = q[sz + sy + 1];
  } else {
    velCorners(c, V, comp, bx, by, bz, cv);
  }
  return trilinear(cv, fx, fy, fz);
}
// windowVel, also returning the range of the eight cells it weighs: [0, 0] when the first
// cell's brick has no slot (the donor bounds of fluid-velocity-transport.wgsl:104-112); for
// cells inside the window that brick's slot is read from the home record.
__device__ __forceinline__ float windowVelBounds(const Ctx &c, const VelWindow &w, const float *V, int comp,
                                                 float px, float py, float pz, float &lo, float &hi) {
  float qx = px - 0.5f, qy = py - 0.5f, qz = pz - 0.5f;
  float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
  int bx = (int)flx, by = (int)fly, bz = (int)flz;
  float fx = qx - flx, fy = qy - fly, fz = qz - flz;
  float cv[8];
  bool none;
  const int rx = bx - w.o[0], ry = by - w.o[1], rz = bz - w.o[2];
  if ((unsigned)rx < (unsigned)(w.e[0] - 1) && (unsigned)ry < (unsigned)(w.e[1] - 1) &&
      (unsigned)rz < (unsigned)(w.e[2] - 1)) {
    const float *q = w.v + __mul24(comp, w.n) + __mul24(rz, w.sz) + __mul24(ry, w.sy) + rx;
    const int sy = w.sy, sz = w.sz, sV = c.L.sV;
    cv[0] = q[0];
    cv[1] = q[1];
    cv[2] = q[sy];
    cv[3] = q[sy + 1];
    cv[4] = q[sz];
    cv[5] = q[sz + 1];
    cv[6] = q[sz + sy];
    cv[7] = q[sz + sy + 1];
    none = c.home.slot[((bz >> sV) - c.home.bz + 1) << 4 | ((by >> sV) - c.home.by + 1) << 2 |
                       ((bx >> sV) - c.home.bx + 1)] < 0;
  } else {
    int slot[8], local[8];
    cornerCells(c, c.L.sV, bx, by, bz, slot, local);
#pragma unroll
    for (int k = 0; k < 8; k++) cv[k] = slot[k] < 0 ? 0.f : V[slotBase(slot[k] * 3 + comp, c.L.sV) + local[k]];
    none = slot[0] < 0;
  }
  const float out = trilinear(cv, fx, fy, fz);
  if (none) {
    lo = hi = 0.f;
  } else {
    lo = 1e5f;
    hi = -1e5f;
    range8(cv, lo, hi);
  }
  return out;
}
// velocityOnFace (fluid-advection.wgsl:26-44) through the window.
__device__ __forceinline__ void velocityOnFace(const Ctx &c, const VelWindow &w, const float *V, const int id[3],
                                               int comp, const float orig[3], float out[3]) {
  float fp[3];
  for (int a = 0; a < 3; a++) fp[a] = (float)id[a] + 0.5f;
  fp[comp] += 0.5f;
  for (int k = 0; k < 3; k++) {
    if (k == comp) {
      out[k] = orig[k];
      continue;
    }
    float pos[3] = {fp[0], fp[1], fp[2]};
    pos[k] = fp[k] - 0.5f;
    out[k] = windowVel(c, w, V, k, pos[0], pos[1], pos[2]);
  }
}
// velocityAt through the window.
__device__ __forceinline__ void velocityAt(const Ctx &c, const VelWindow &w, const Params &P, const float *V,
                                           float wx, float wy, float wz, float out[3]) {
  const float cx = wx * P.invHv, cy = wy * P.invHv, cz = wz * P.invHv;
  out[0] = windowVel(c, w, V, 0, (wx - 0.5f * P.hv) * P.invHv, cy, cz);
  out[1] = windowVel(c, w, V, 1, cx, (wy - 0.5f * P.hv) * P.invHv, cz);
  out[2] = windowVel(c, w, V, 2, cx, cy, (wz - 0.5f * P.hv) * P.invHv);
}

}  // namespace fp
