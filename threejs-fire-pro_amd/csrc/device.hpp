// Device-side building blocks of the solver: parameters, directory lookups, trilinear
// sampling and scalar math. Every float expression is written in the evaluation order the
// kernels' contract fixes (no contraction: the library is built with -ffp-contract=off),
// so a result does not depend on the compiler's scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fp {

constexpr int FRESH_ENTRY = 0x2000000;        // tiles.wgsl:24
constexpr uint32_t FRESH_SLOT = 0x80000000u;  // fluid-allocation.wgsl:30
constexpr uint32_t CONTENT = 1u, EXPANDS = 2u, MAY_EXPAND = 4u;
constexpr uint32_t BRICK_FREES = 8u, BRICK_REQUESTS = 16u;
constexpr int RECORD_WORDS = 28, PRESSURE_FLAG = 27;
constexpr int TILE_RECORD_WORDS = 32;
constexpr uint8_t STORED_SOLID = 64, UNKNOWN = 128;
constexpr int MAX_EMITTERS = 64, MAX_FORCES = 16, MAX_COLLIDERS = 16, MAX_REGIONS = 64;

// Scene records as the kernels read them.
struct Emitter {
  float pos[3], size[3];
  int box;
  float flame, heat, smoke, response;
  float vel[3];
  float outward, period, strength, fuel;
  float invSize[3];  // 1 / max(size, 0.03)
  float invPeriod;   // 1 / period
};
struct Collider {
  float center[3], extent[3];
  int box;           // 0 sphere, 1 box, 2 mesh occupancy
  float axes[3][3];  // columns
  int lo[3], size[3], maskOffset;  // mesh: its box of field cells and its first mask word
};
struct Region {
  float center[3], kind, size[3];
  float invSize[3];  // 1 / max(size, 0.03)
};
struct ForceRec {
  float axis[3], strength, center[3], kind, scale, lift, inward;
  int targeted, firstRegion, regionCount;
  int meshTargets;  // its targets include mesh emitters: their covered velocity cells count 1
};

// Per-step parameters (fluid-common.wgsl Params + FlameParams), by value.
struct Params {
  float hf, hv, dt;
  float invHf, invHv;  // 1/hf and 1/hv, rounded once on the host
  float buoyancy, cooling, smokeWeight, damping, dissipation, vorticity, cutoff;
  int lifetimeZero;
  float time, seed;
  uint32_t seedBits;
  int colliders, ground;
  float ignition, burnRate;
  int fuelEnabled;
  float fdt, lifespan, fsmoke, fheat, fexpansion;
  int knotCount[3];
  float knots[3][8][2];
  int emitterCount, forceCount;
  int maccormack;
  // Step-uniform factors, evaluated once on the host with the kernels' own fexp:
  // e^{-burnRate dt}, e^{-dissipation dt}, e^{-cooling dt}, the same over dt / 2,
  // e^{-damping dt}, and the flame's lifetime decrement fdt / lifespan.
  float burnKeep, lossS, lossH, halfS, halfH, damp, lifeStep;
  // Coarse smoke (smokeDivisor 2 or 4): the smoke grid's spacing and its reciprocal, and
  // the smoke cells a velocity cell spans on each axis (1 or 2).
  int coarse;
  float hs, invHs;
  int smokeRatio;
  int meshRecords;  // mesh emission records this step (0: no mesh sources)
  // A box of field cells (inclusive) holding every cell any collider makes solid: a solid
  // bit outside it is never set, so lookups outside it are skipped.
  int solidLo[3], solidHi[3];
};
// Whether field cell (x, y, z) lies in the colliders' box (Params::solidLo/solidHi).
__device__ __forceinline__ bool inSolidBox(const Params &P, int x, int y, int z) {
  return x >= P.solidLo[0] && x <= P.solidHi[0] && y >= P.solidLo[1] && y <= P.solidHi[1] && z >= P.solidLo[2] &&
         z <= P.solidHi[2];
}

// Geometry of the grids and pools (constant per simulation).
struct Layout {
  int B, b, sF, sV, nF, nV;  // field and velocity brick edges, shifts, cells
  int side;                   // pool slots per row and rows per layer
  int brickLevels;
  int sd, bs, sS, nS;         // smoke divisor, the smoke brick's edge, shift and cells
};

// The tile directory and brick entries as device arrays.
struct Dir {
  const int *tileRec;  // [cap * 32]
  const int *hash;     // [size * 4]
  uint32_t hashMask;
  const int *entries;  // [cap * 64]
  const int *records;  // [cap * 64 * 28]
  int capacity;
};

__device__ __forceinline__ uint32_t tileHash(int x, int y, int z) {
  return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}

// findTile (tiles.wgsl:57-73): open addressing, at most half full.
__device__ __forceinline__ int findTile(const Dir &d, int x, int y, int z) {
  uint32_t h = tileHash(x, y, z) & d.hashMask;
  for (uint32_t probe = 0; probe <= d.hashMask; probe++) {
    const int4 e = *reinterpret_cast<const int4 *>(d.hash + (size_t)h * 4);
    if (e.w < 0) return -1;
    if (e.x == x && e.y == y && e.z == z) return e.w;
    h = (h + 1u) & d.hashMask;
  }
  return -1;
}

__device__ __forceinline__ int brickIdOf(const Dir &d, int bx, int by, int bz) {
  int t = findTile(d, bx >> 2, by >> 2, bz >> 2);
  if (t < 0) return -1;
  return t * 64 + (((bz & 3) * 4 + (by & 3)) * 4 + (bx & 3));
}

__device__ __forceinline__ int entryLookup(const Dir &d, int bx, int by, int bz) {
  int id = brickIdOf(d, bx, by, bz);
  return id < 0 ? -1 : d.entries[id];
}

__device__ __forceinline__ int pageToSlot(int page, int side) {
  uint32_t p = (uint32_t)page;
  // A page holds 8-bit coordinates, so side <= 256 and the masks keep every product in the
  // range of the full-rate 24-bit multiplies.
  const uint32_t r = (uint32_t)side & 511u;
  return (int)((((p >> 16) & 255u) * r + ((p >> 8) & 255u)) * r + (p & 255u));
}
__device__ __forceinline__ int slotToPage(int slot, int side) {
  uint32_t s = (uint32_t)slot, r = (uint32_t)side;
  return (int)((s % r) | (((s / r) % r) << 8) | ((s / (r * r)) << 16));
}
__device__ __forceinline__ int entrySlot(int e, int side) {
  return e < 0 ? -1 : pageToSlot(e & 0xffffff, side);
}

// A work item's home brick: its coordinates and the 27 neighbor entries of its record,
// staged in LDS, so lookups within one brick need no hash probe. The entries are in (z * 3 +
// y) * 3 + x order (neighbor offsets plus one); the same bricks' pool slots sit at (z << 4) |
// (y << 2) | x, so that the hot slot lookups index with shifts, not quarter-rate multiplies.
constexpr int HOME_SLOT_WORDS = 43;
constexpr int HOME_WORDS = 27 + HOME_SLOT_WORDS;
__device__ __forceinline__ int homeSlotWord(int w) { return (w / 9) << 4 | ((w / 3) % 3) << 2 | (w % 3); }
struct Home {
  int bx, by, bz;
  const int *rec;   // 27 entries (LDS)
  const int *slot;  // the same 27 bricks' pool slots, -1 without one (LDS, shift layout)
};

__device__ __forceinline__ int entryAt(const Home &h, const Dir &d, int bx, int by, int bz) {
  int dx = bx - h.bx, dy = by - h.by, dz = bz - h.bz;
  if ((unsigned)(dx + 1) <= 2u && (unsigned)(dy + 1) <= 2u && (unsigned)(dz + 1) <= 2u)
    return h.rec[(dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
  return entryLookup(d, bx, by, bz);
}

// Whether every brick of the box [lo, hi] (at most two bricks a side) has an entry, as loops
// over the box would find: the box's eight corners tested without loops or early exits
// (duplicates where lo == hi), so lanes do not diverge on trip counts. A box empty on an axis
// holds no brick: true.
// NEAR: the box lies in the home brick's 3^3 neighborhood, read from the LDS record alone.
template <bool NEAR = false>
__device__ __forceinline__ bool bricksPresent(const Home &h, const Dir &d, const int lo[3], const int hi[3]) {
  if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return true;
  bool all = true;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int b[3] = {(k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 4) ? hi[2] : lo[2]};
    const int e = NEAR ? h.rec[(b[2] - h.bz + 1) * 9 + (b[1] - h.by + 1) * 3 + (b[0] - h.bx + 1)]
                       : entryAt(h, d, b[0], b[1], b[2]);
    all &= e >= 0;
  }
  return all;
}

// The work item of (virtual) workgroup `w` among `count`, such that each XCD works through one
// contiguous run of items. MI355X hands workgroups to its 8 XCDs round-robin by id, and
// each XCD has its own L2; items in brick-id order are spatial neighbors, so a contiguous
// run per XCD keeps the neighbor reads of a brick in the L2 that already holds them.
// Launches must have at least 8 * ceil(count / 8) workgroups. -1 when there is no item.
constexpr int XCDS = 8;
__device__ __forceinline__ int xcdItem(int w, int count) {
  const int per = (count + XCDS - 1) / XCDS;
  const int j = w / XCDS;
  if (j >= per) return -1;
  const int i = (w % XCDS) * per + j;
  return i < count ? i : -1;
}

// The OR of v over the wavefront's lanes, in every lane (all lanes active).
__device__ __forceinline__ uint32_t waveOr(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
  return v;
}

// ---- scalar math, identical trees on host and device -----------------------------------
__device__ __forceinline__ float fminv(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmaxv(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float fabsv(float a) { return a < 0.0f ? -a : a; }
// The least and greatest of three, one VOP3 instruction for two comparisons where the
// select form above takes a compare and a select per value (and the library fminf a
// canonicalizing move per loaded operand). For operands that are not NaN the result is the
// value a chain of fminv/fmaxv gives, except that a zero extremum may carry the other sign
// when the operands hold both +0 and -0: used for donor ranges only, which enter comparisons
// (sign-blind) and, for a range of one value, the scalar fields, which hold no -0.
__device__ __forceinline__ float min3v(float a, float b, float c) {
  float r;
  asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
__device__ __forceinline__ float max3v(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
// The range of eight values from the starting bounds lo and hi.
__device__ __forceinline__ void range8(const float v[8], float &lo, float &hi) {
  lo = min3v(min3v(min3v(min3v(lo, v[0], v[1]), v[2], v[3]), v[4], v[5]), v[6], v[7]);
  hi = max3v(max3v(max3v(max3v(hi, v[0], v[1]), v[2], v[3]), v[4], v[5]), v[6], v[7]);
}
__device__ __forceinline__ float clampv(float x, float lo, float hi) { return fminv(fmaxv(x, lo), hi); }
__device__ __forceinline__ float mixv(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float smoothstepv(float e0, float e1, float x) {
  float t = clampv((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
// sqrtf is the correctly rounded square root (v_sqrt_f32 plus its fix-up); __fsqrt_rn
// compiles to the bare approximation on gfx950.
__device__ __forceinline__ float length3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// exp in single precision: Cody & Waite reduction x = k ln2 + r with a two-part ln2,
// the degree-7 Taylor polynomial of e^r (|r| <= ln2/2) in Horner form, and an exact
// scaling by 2^k. Relative error below 2 ulp; evaluated with IEEE basic operations only.
__host__ __device__ __forceinline__ float bitsFloat(int i) { return __builtin_bit_cast(float, i); }
__host__ __device__ __forceinline__ float fexp(float x) {
  if (!(x < 88.7228f)) return x != x ? x : bitsFloat(0x7f800000);
  if (x < -103.972f) return 0.0f;
  float k = floorf(x * 1.44269502f + 0.5f);
  float r = x - k * 0.693145752f;
  r = r - k * 1.42860677e-06f;
  float p = 1.98412701e-04f;
  p = p * r + 1.38888892e-03f;
  p = p * r + 8.33333377e-03f;
  p = p * r + 4.16666679e-02f;
  p = p * r + 1.66666672e-01f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  int ki = (int)k;
  if (ki > 127) {
    p = p * 2.0f;
    ki -= 1;
  }
  if (ki > -126) return p * bitsFloat((ki + 127) << 23);
  return (p * bitsFloat((ki + 64 + 127) << 23)) * bitsFloat((127 - 64) << 23);
}

// 2^x in single precision for -126 <= x <= 127: k = the nearest integer, r = (x - k) ln2
// (x - k is exact, |r| <= ln2/2), the polynomial of fexp for e^r, and an exact 2^k.
__device__ __forceinline__ float fexp2(float x) {
  float k = floorf(x + 0.5f);
  float r = (x - k) * 0.693147182f;
  float p = 1.98412701e-04f;
  p = p * r + 1.38888892e-03f;
  p = p * r + 8.33333377e-03f;
  p = p * r + 4.16666679e-02f;
  p = p * r + 1.66666672e-01f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return p * __int_as_float(((int)k + 127) << 23);
}

}  // namespace fp
