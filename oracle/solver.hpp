// ORACLE — test infrastructure only. Nothing in the shipped library may include, link or
// call this file; tests, smoke() and bench.py's CPU baseline use it as the checker.
//
// A sequential CPU restatement of one Fire Pro solver step (src/engine/FluidSimulation.ts
// substep, :1374-1531) and its WGSL kernels, templated on the arithmetic type: float for
// the bit-exact twin of the HIP path, double for the FP64 error reference.
//
// It restates the reference's *logical* semantics:
//  * a brick's cells hold values while it has a slot; bricks without one read zero
//    (fluid-sampling.wgsl:217-225). The reference's texture aprons only make hardware
//    filtering see those logical values (fluid-sampling.wgsl:232-264 reads cell by cell
//    exactly where an apron could be stale), so samples here are exact trilinear sums;
//  * per-tile scene lists (FluidSimulation.ts:1150-1366) are culling: each list holds every
//    emitter, force and collider that can reach the tile, so all are visited in index order;
//  * slots are freed and granted in brick-id order (fluid-allocation.wgsl:246-250), so the
//    allocator is a sequential loop; lists that the reference appends with atomics are
//    produced in brick-id order (their sets are what the reference fixes).
// Deliberate differences from the reference, which stores rgba16float textures and f16
// buffers: values are stored at the template precision, with no half-precision rounding
// (truncateHalf, fluid-common.wgsl:107-114, and every f16 store); the activity readback is
// applied before the next step (the earliest a mapAsync can resolve,
// FluidSimulation.ts:1701-1746); the ambient-free debug exports are not restated.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../include/firepro.h"
#include "directory.hpp"
#include "fmath.hpp"

namespace oracle {

constexpr int FRESH_ENTRY = 0x2000000;          // tiles.wgsl:24
constexpr uint32_t FRESH_SLOT = 0x80000000u;    // fluid-allocation.wgsl:30
constexpr uint32_t CONTENT = 1, EXPANDS = 2, MAY_EXPAND = 4;  // fluid-content.wgsl:11-13
constexpr uint32_t BRICK_FREES = 8, BRICK_REQUESTS = 16;      // fluid-allocation.wgsl:38-40
constexpr int MIN_TILES = 16;                   // FluidSimulation.ts:69
constexpr int BRICK_RECORD_WORDS = 28;          // FluidSimulation.ts:137
constexpr int PRESSURE_FLAG = 27;               // tiles.wgsl:21
constexpr int BOTTOM_ITERATIONS = 19;           // PressureSolver.ts:326
constexpr int STORED_SOLID = 64;                // pressure.wgsl:170
constexpr double MAX_DT = 1.0 / 30.0;

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline int ilog2(int v) {
  int s = 0;
  while ((1 << s) < v) s++;
  return s;
}

// Packed scene records (fluid-common.wgsl:34-51, force-fields.wgsl:4-26).
template <class R> struct Emitter {
  R pos[3], size[3];
  bool box;
  R flame, heat, smoke, response;
  R vel[3];
  R outward, period, strength, fuel;
  R invSize[3];  // 1 / max(size, 0.03)
  R invPeriod;   // 1 / period
};
template <class R> struct Collider {
  R center[3], extent[3];
  int box;       // 0 sphere, 1 box, 2 mesh occupancy
  R axes[3][3];  // columns a, b, c
  int lo[3] = {0, 0, 0}, size[3] = {0, 0, 0};
  std::vector<uint8_t> inside;  // mesh: occupancy of its box of cells, x fastest
};
template <class R> struct Region {
  R center[3];
  R kind;  // 1 sphere, 3 box
  R size[3];
  R invSize[3];  // 1 / max(size, 0.03)
};
template <class R> struct Force {
  R axis[3], strength, center[3], kind, scale, lift, inward;
  bool targeted;
  std::vector<Region<R>> regions;
  bool meshTargets = false;
  std::map<std::array<int, 3>, bool> covered;  // velocity cells its mesh targets cover
};

// World-to-cell conversions multiply by the reciprocal of the spacing, rounded once: WGSL
// allows its division 2.5 ulp, and the product stays within that.
template <class R> struct Params {
  R hf, hv, dt;
  R invHf, invHv;
  // Coarse smoke (smokeDivisor > 1): the smoke grid's spacing and its reciprocal.
  R hs, invHs;
  R buoyancy, cooling, smokeWeight, damping, dissipation, vorticity, cutoff;
  bool lifetimeZero;
  R time, seed;
  uint32_t seedBits;
  int colliders;
  bool ground;
  R ignition, burnRate;
  bool fuelEnabled;
  // Flame step (flame.ts:43-87): dt, lifespan, smoke/s, heat/s, expansion rate, ramps.
  R fdt, lifespan, fsmoke, fheat, fexpansion;
  int knotCount[3];
  R knots[3][8][2];
};

// Sum of a piecewise-linear ramp's area from lo to hi (flame-common.wgsl:14-28).
template <class R> R rampIntegral(const Params<R> &p, int ch, R lo, R hi) {
  R area = 0;
  for (int i = 1; i < p.knotCount[ch]; i++) {
    R ax = p.knots[ch][i - 1][0], ay = p.knots[ch][i - 1][1];
    R bx = p.knots[ch][i][0], by = p.knots[ch][i][1];
    R left = fmax(lo, ax), right = fmin(hi, bx);
    if (right > left) {
      R y0 = mixv(ay, by, (left - ax) / (bx - ax));
      R y1 = mixv(ay, by, (right - ax) / (bx - ax));
      area += (right - left) * (y0 + y1) * R(0.5);
    }
  }
  return area;
}

// evaluateFlame (flame-common.wgsl:30-42): lifetime after the step, and smoke, heat and
// expansion produced over it.
template <class R> void evaluateFlame(const Params<R> &p, R before, R &after, R prod[3]) {
  if (before <= R(0)) {
    after = 0;
    prod[0] = prod[1] = prod[2] = 0;
    return;
  }
  after = fmax(R(0), before - p.fdt / p.lifespan);
  prod[0] = p.fsmoke * p.lifespan * rampIntegral(p, 0, after, before);
  prod[1] = p.fheat * p.lifespan * rampIntegral(p, 1, after, before);
  prod[2] = p.fexpansion * (p.lifespan * rampIntegral(p, 2, after, before) / p.fdt);
}

// emitterWeight (fluid-sources.wgsl:24-38).
template <class R> R emitterWeight(const Emitter<R> &e, const R w[3]) {
  // Past the padded size on any axis |d| >= 1 there, so r >= 1: no weight.
  for (int a = 0; a < 3; a++)
    if (fabs_(w[a] - e.pos[a]) >= fmax(e.size[a], R(0.03))) return 0;
  R d[3];
  for (int a = 0; a < 3; a++) d[a] = (w[a] - e.pos[a]) * e.invSize[a];
  R r = e.box ? fmax(fabs_(d[0]), fmax(fabs_(d[1]), fabs_(d[2]))) : length3(d[0], d[1], d[2]);
  if (r >= R(1)) return 0;
  return R(1) - smoothstepv(R(0.5), R(1), r);
}

// burstGradient / burstPerlin / burstPattern / burstVariation (fluid-sources.wgsl:40-82).
template <class R> R burstGradient(uint32_t seedBits, int cx, int cy, int cz, R ox, R oy, R oz) {
  uint32_t h = (uint32_t)cx * 374761393u + (uint32_t)cy * 668265263u + (uint32_t)cz * 2246822519u +
               seedBits * 3266489917u;
  h = (h ^ (h >> 13)) * 1274126177u;
  h = (h ^ (h >> 16)) & 15u;
  R a = h < 8u ? ox : oy;
  R b = h < 4u ? oy : ((h == 12u || h == 14u) ? ox : oz);
  return ((h & 1u) == 0u ? a : -a) + ((h & 2u) == 0u ? b : -b);
}
template <class R> R burstPerlin(uint32_t seedBits, R px, R py, R pz) {
  R flx = ffloor(px), fly = ffloor(py), flz = ffloor(pz);
  int cx = (int)flx, cy = (int)fly, cz = (int)flz;
  R fx = px - flx, fy = py - fly, fz = pz - flz;
  auto fade = [](R f) { return f * f * f * (f * (f * R(6) - R(15)) + R(10)); };
  R sx = fade(fx), sy = fade(fy), sz = fade(fz);
  R a = burstGradient<R>(seedBits, cx, cy, cz, fx, fy, fz);
  R b = burstGradient<R>(seedBits, cx + 1, cy, cz, fx - 1, fy, fz);
  R c = burstGradient<R>(seedBits, cx, cy + 1, cz, fx, fy - 1, fz);
  R d = burstGradient<R>(seedBits, cx + 1, cy + 1, cz, fx - 1, fy - 1, fz);
  R e = burstGradient<R>(seedBits, cx, cy, cz + 1, fx, fy, fz - 1);
  R g = burstGradient<R>(seedBits, cx + 1, cy, cz + 1, fx - 1, fy, fz - 1);
  R h = burstGradient<R>(seedBits, cx, cy + 1, cz + 1, fx, fy - 1, fz - 1);
  R j = burstGradient<R>(seedBits, cx + 1, cy + 1, cz + 1, fx - 1, fy - 1, fz - 1);
  return mixv(mixv(mixv(a, b, sx), mixv(c, d, sx), sy), mixv(mixv(e, g, sx), mixv(h, j, sx), sy), sz);
}
template <class R> R burstPattern(uint32_t seedBits, R px, R py, R pz) {
  R n = (burstPerlin<R>(seedBits, px, py, pz) +
         R(0.35) * burstPerlin<R>(seedBits, px * R(2) + R(5.2), py * R(2) + R(9.7), pz * R(2) + R(3.1))) /
        R(1.35);
  return clampv(R(0.5) + R(0.5) * n, R(0), R(1));
}
template <class R> void burstVariation(uint32_t seedBits, const Emitter<R> &e, const R w[3], R out[2]) {
  if (e.strength == R(0)) {
    out[0] = out[1] = 1;
    return;
  }
  R q[3];
  const R off[3] = {R(3.17), R(7.53), R(11.29)};
  for (int a = 0; a < 3; a++) q[a] = (w[a] - e.pos[a]) * e.invPeriod + off[a];
  R s0 = burstPattern<R>(seedBits, q[0], q[1], q[2]);
  R s1 = burstPattern<R>(seedBits, q[0] + R(17.3), q[1] + R(2.8), q[2] + R(8.6));
  out[0] = R(1) - e.strength * s0;
  out[1] = R(1) - e.strength * s1;
}

// curl-noise.wgsl:1-45: the periodic 64^3 curl-noise table, built once.
inline float noiseHash(float px, float py, float pz) {
  uint32_t v[3] = {(uint32_t)(int)px & 15u, (uint32_t)(int)py & 15u, (uint32_t)(int)pz & 15u};
  for (auto &c : v) c = c * 1664525u + 1013904223u;
  v[0] += v[1] * v[2];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
  for (auto &c : v) c ^= c >> 16;
  v[0] += v[1] * v[2];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
  return (float)(v[0] >> 8) / 16777216.0f;
}
inline void noiseGradient(float px, float py, float pz, float out[3]) {
  float ix = std::floor(px), iy = std::floor(py), iz = std::floor(pz);
  float fx = px - ix, fy = py - iy, fz = pz - iz;
  auto s = [](float f) { return f * f * f * (f * (f * 6.0f - 15.0f) + 10.0f); };
  auto ds = [](float f) { return 30.0f * f * f * (f - 1.0f) * (f - 1.0f); };
  float sx = s(fx), sy = s(fy), sz = s(fz);
  float a = noiseHash(ix, iy, iz), b = noiseHash(ix + 1, iy, iz), c = noiseHash(ix, iy + 1, iz);
  float d = noiseHash(ix + 1, iy + 1, iz), e = noiseHash(ix, iy, iz + 1);
  float f1 = noiseHash(ix + 1, iy, iz + 1), g = noiseHash(ix, iy + 1, iz + 1);
  float h = noiseHash(ix + 1, iy + 1, iz + 1);
  out[0] = mixv(mixv(b - a, d - c, sy), mixv(f1 - e, h - g, sy), sz) * ds(fx);
  out[1] = mixv(mixv(c - a, d - b, sx), mixv(g - e, h - f1, sx), sz) * ds(fy);
  out[2] = mixv(mixv(e - a, f1 - b, sx), mixv(g - c, h - d, sx), sy) * ds(fz);
}
inline std::vector<float> buildCurlNoise() {
  std::vector<float> t(64 * 64 * 64 * 4, 0.0f);
  for (int z = 0; z < 64; z++)
    for (int y = 0; y < 64; y++)
      for (int x = 0; x < 64; x++) {
        float p[3] = {((float)x + 0.5f) / 64.0f * 16.0f, ((float)y + 0.5f) / 64.0f * 16.0f,
                      ((float)z + 0.5f) / 64.0f * 16.0f};
        float dx[3], dy[3], dz[3];
        noiseGradient(p[0], p[1], p[2], dx);
        noiseGradient(p[0] + 7.1f, p[1] + 3.7f, p[2] + 11.3f, dy);
        noiseGradient(p[0] + 5.2f, p[1] + 7.9f, p[2] + 9.5f, dz);
        float *o = &t[(((size_t)z * 64 + y) * 64 + x) * 4];
        o[0] = dz[1] - dy[2];
        o[1] = dx[2] - dz[0];
        o[2] = dy[0] - dx[1];
        o[3] = 0;
      }
  return t;
}

template <class R> class Sim {
 public:
  // Construction (FluidSimulation.ts:359-418).
  explicit Sim(const fp_grid &g) {
    if (!std::isfinite(g.voxel_size) || g.voxel_size < 0.0001 || g.voxel_size > 16)
      throw Error("voxelSize must be from 0.0001 to 16 meters.");
    voxel = g.voxel_size;
    d = g.velocity_divisor;
    if (d != 1 && d != 2 && d != 4) throw Error("velocityDivisor must be 1, 2 or 4.");
    B = g.brick_size;
    if (B != 8 && B != 16 && B != 32) throw Error("brickSize must be 8, 16 or 32 fine cells.");
    smokeDiv = g.smoke_divisor;
    if (smokeDiv != 1 && smokeDiv != 2 && smokeDiv != 4) throw Error("smokeDivisor must be 1, 2 or 4.");
    maccormack = g.scalar_maccormack != 0;
    b = B / d;
    sF = ilog2(B);
    sV = ilog2(b);
    nF = B * B * B;
    nV = b * b * b;
    bs = B / smokeDiv;
    sS = ilog2(bs);
    nS = bs * bs * bs;
    smokeRatio = d > smokeDiv ? d / smokeDiv : 1;
    side = poolSide(d, B);
    maxSlots = side * side * poolSlotsPerAxis(B);
    brickLevels = sV;
    bottom = brickLevels + 2;
    noise = buildCurlNoise();
    allocate();
  }

  // ---- host state -------------------------------------------------------------------
  double voxel;
  int d, B, b, sF, sV, nF, nV, side, maxSlots, brickLevels, bottom;
  // Coarse smoke: the divisor, the smoke brick's edge, shift and cells, and the smoke
  // cells per velocity cell edge that buoyancy averages (FluidSimulation.ts:374-380).
  int smokeDiv = 1, bs = 0, sS = 0, nS = 0, smokeRatio = 1;
  bool maccormack = true;
  double voxelBudget = INFINITY;
  double cutoff = 0;
  bool ground = true;
  int poolSlots = 0, usedSlots = 0, poolTarget = 0;
  int64_t poolGrownAt = 0;
  bool poolLimited = false;
  int64_t activeBrickCount = 0;
  double time = 0;
  int64_t steps = 0;
  bool mayHaveFlame = false;
  bool solidsStale = true, geometryDirty = true;
  int colliderCount = 0;
  std::vector<uint32_t> prevColliders;
  std::vector<uint8_t> prevOccupancy;
  std::vector<uint64_t> sourceIds;
  std::vector<std::pair<std::array<int, 3>, uint32_t>> contentTiles;
  int64_t contentStep = -1;
  Directory dir;
  // The boxes the renderer marches (FluidSimulation.ts:163-174, :296-305): lo and hi
  // bricks per cluster with something to draw, the clusters they box, every tile slot's
  // coordinates with its box index (or -1), and a version that changes with either.
  struct BrickBox {
    int lo[3], hi[3];
  };
  std::vector<BrickBox> renderBoxes;
  std::vector<int> renderClusters;
  std::vector<int32_t> renderTiles;
  int renderVersion = 0;
  bool renderStale = true;
  int capacity = 0;
  std::vector<float> noise;

  // Per brick id (tile resources, FluidSimulation.ts:139-160).
  std::vector<int32_t> pages, entries, records;
  std::vector<uint32_t> brickState, activity;
  std::vector<uint32_t> freeSlots;
  int freeCount = 0;
  uint32_t failures = 0;
  std::vector<int32_t> activeList;
  std::vector<std::pair<int32_t, uint32_t>> zeroList;
  uint32_t lastFailures = 0;

  // Pools, slot-major: field cells x4, expansion, velocity [slot][axis][cell], solids bits.
  std::vector<R> F, Fp, Fn, E, V, Vadv, Vcor, Vf, Curl, D, Pfine;
  // The last step's MacCormack limiter decisions per field cell, for the precision report:
  // bits 0-3 the corrected smoke, heat, lifetime and fuel kept (inside the donors' range),
  // bit 4 the smoke/heat/fuel limiter evaluated, bit 5 the lifetime one (0: first order, a
  // solid cell or donors all equal). Each cell writes its own byte.
  mutable std::vector<uint8_t> limiter;
  std::vector<uint32_t> solids;
  // Coarse smoke pools (smoke.wgsl): density, predictor, next, and the predictor's donor
  // bounds per smoke cell.
  std::vector<R> S, Sp, Sn, SLo, SHi;
  // Multigrid levels: brick levels per slot, tile levels per tile slot.
  struct Level {
    int side = 0;
    bool tile = false;
    std::vector<R> p, q, rhs;
    std::vector<uint8_t> topo;
  };
  std::vector<Level> levels;
  // Pressure policy (no counterpart in the reference): 0 one cold V-cycle; 1, 2, 3 defect
  // correction with V-cycle levels stored in FP32, BF16, FP16 (storage rounding applies to
  // the float restatement) and the coarsest level in FP64.
  int pressureMode = 0;
  double pressureTol = 1e-6;
  int pressureMaxCycles = 20;
  int lastCycles = 1;
  double lastResidual = -1;
  std::vector<R> dcX, dcXlo;  // the solution as an unevaluated sum hi + lo
  std::vector<R> PfineLo;     // the pressure's low part when it is such a pair, else empty
  R stored(R v) const {
    if constexpr (std::is_same<R, float>::value) {
      if (pressureMode == 2) return roundBF16(v);
      if (pressureMode == 3) return roundF16(v);
    }
    return v;
  }

  // Dense directory view for cell lookups: each brick's entry in the tiles' bounding box.
  int glo[3] = {0, 0, 0}, gdim[3] = {0, 0, 0};
  std::vector<int32_t> egrid;

  Params<R> P;
  std::vector<Emitter<R>> emitters;
  // Mesh emission: records of 8 values behind a zero record 0 (flame, heat, smoke, fuel,
  // motion xyz, response), stored as f32 like the reference's record buffer, and the brick
  // tables (per brick id a table or -1, then per table a record per field cell).
  std::vector<R> meshRecords{std::vector<R>(8, R(0))};
  std::vector<int32_t> meshTables;
  int meshCount = 0;
  std::vector<Collider<R>> colliders;
  std::vector<Force<R>> forces;

  int brickCount() const { return capacity * 64; }
  int slotLimit() const {
    double voxels = std::min(voxelBudget, maxPoolVoxels(d, B));
    return std::max(1, (int)std::min((double)maxSlots, std::floor(voxels / std::pow((double)B, 3))));
  }
  int pageSlot(int page) const {
    uint32_t p = (uint32_t)page;
    return (int)(((p >> 16) * side + ((p >> 8) & 255u)) * side + (p & 255u));
  }
  int slotPage(int slot) const {
    uint32_t s = (uint32_t)slot, row = (uint32_t)side;
    return (int)((s % row) | (((s / row) % row) << 8) | ((s / (row * row)) << 16));
  }

  // ---- allocation of pools and tiles (FluidSimulation.ts:487-532) ---------------------
  void allocate() {
    int slots = initialPoolSlots(side, slotLimit());
    poolSlots = slots;
    usedSlots = 0;
    poolTarget = 0;
    sourceIds.clear();
    poolLimited = false;
    resizePools(slots, false);
    freeSlots.assign(slots, 0);
    for (int i = 0; i < slots; i++) freeSlots[i] = (uint32_t)(slots - 1 - i);
    freeCount = slots;
    failures = 0;
    solidsStale = true;
    geometryDirty = true;
    dir.clear();
    contentTiles.clear();
    contentStep = -1;
    capacity = 0;
    createTiles(MIN_TILES);
    applyTiles();
    updateBoxes({});
  }

  void resizePools(int slots, bool keep) {
    auto rs = [&](std::vector<R> &v, size_t per) {
      if (keep) v.resize((size_t)slots * per, R(0));
      else v.assign((size_t)slots * per, R(0));
    };
    rs(F, (size_t)nF * 4);
    rs(E, (size_t)nF);
    rs(V, (size_t)nV * 3);
    if (coarse()) {
      rs(S, (size_t)nS);
      for (auto *v : {&Sp, &Sn, &SLo, &SHi}) v->assign((size_t)slots * nS, R(0));
    }
    Fp.assign((size_t)slots * nF * 4, R(0));
    Fn.assign((size_t)slots * nF * 4, R(0));
    limiter.assign((size_t)slots * nF, 0);
    Vadv.assign((size_t)slots * nV * 3, R(0));
    Vcor.assign((size_t)slots * nV * 3, R(0));
    Vf.assign((size_t)slots * nV * 3, R(0));
    Curl.assign((size_t)slots * nV * 4, R(0));
    D.assign((size_t)slots * nV, R(0));
    Pfine.assign((size_t)slots * nV, R(0));
    PfineLo.clear();
    if (keep) solids.resize((size_t)slots * nF / 32, 0u);
    else solids.assign((size_t)slots * nF / 32, 0u);
    levels.assign(bottom + 1, Level());
    for (int L = 0; L < brickLevels; L++) {
      levels[L].side = b >> L;
      levels[L].tile = false;
    }
    for (int L = brickLevels; L <= bottom; L++) {
      levels[L].side = 4 >> (L - brickLevels);
      levels[L].tile = true;
    }
    sizeLevels(slots);
  }
  void sizeLevels(int slots) {
    for (int L = 0; L <= bottom; L++) {
      Level &lv = levels[L];
      size_t owners = lv.tile ? (size_t)std::max(capacity, 1) : (size_t)slots;
      size_t n = owners * lv.side * lv.side * lv.side;
      lv.p.assign(n, R(0));
      lv.q.assign(n, R(0));
      lv.rhs.assign(n, R(0));
      lv.topo.assign(n, 0);
    }
  }

  void createTiles(int cap) {
    int old = capacity;
    size_t bricks = (size_t)cap * 64;
    pages.resize(bricks, -1);
    entries.resize(bricks, -1);
    records.resize(bricks * BRICK_RECORD_WORDS, 0);
    brickState.resize(bricks, 0);
    // Brick flags, then one content word per tile slot. Growth copies only the bricks'.
    std::vector<uint32_t> act(bricks + cap, 0u);
    for (size_t i = 0; i < (size_t)old * 64; i++) act[i] = activity[i];
    activity.swap(act);
    capacity = cap;
    for (int L = brickLevels; L <= bottom; L++) {
      Level &lv = levels[L];
      size_t n = (size_t)cap * lv.side * lv.side * lv.side;
      lv.p.assign(n, R(0));
      lv.q.assign(n, R(0));
      lv.rhs.assign(n, R(0));
      lv.topo.assign(n, 0);
    }
  }

  Tables tables;
  // applyTiles (FluidSimulation.ts:663-737).
  void applyTiles() {
    int needed = std::max(MIN_TILES, dir.slotsUsed());
    if (needed > capacity) {
      int cap = capacity;
      while (cap < needed) cap *= 2;
      createTiles(cap);
    }
    tables = dir.tables(capacity);
    for (int slot : dir.takeCreated()) {
      for (int i = 0; i < 64; i++) {
        size_t id = (size_t)slot * 64 + i;
        pages[id] = -1;
        brickState[id] = 0;
        activity[id] = 0;
        entries[id] = -1;
      }
    }
    geometryDirty = true;  // setBricks resets the solver's plans (PressureSolver.ts:120-134)
    renderStale = true;    // FluidSimulation.ts:708-709
  }

  // ---- per-brick lookups (tiles.wgsl:75-181) -----------------------------------------
  int brickIdOf(int bx, int by, int bz) const {
    int slot = dir.slotAt(bx >> 2, by >> 2, bz >> 2);
    if (slot < 0) return -1;
    return slot * 64 + ((bz & 3) * 4 + (by & 3)) * 4 + (bx & 3);
  }
  void brickOf(int id, int out[3]) const {
    const int32_t *r = &tables.records[(size_t)(id >> 6) * TILE_RECORD_WORDS];
    int l = id & 63;
    out[0] = r[0] * 4 + (l & 3);
    out[1] = r[1] * 4 + ((l >> 2) & 3);
    out[2] = r[2] * 4 + (l >> 4);
  }
  int tileState(int tileSlot) const { return tables.records[(size_t)tileSlot * TILE_RECORD_WORDS + 3]; }

  void buildEntryGrid() {
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (Tile *t : dir.order)
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], t->c[a] * 4);
        hi[a] = std::max(hi[a], t->c[a] * 4 + 3);
      }
    if (dir.order.empty()) {
      for (int a = 0; a < 3; a++) {
        glo[a] = 0;
        gdim[a] = 0;
      }
      egrid.clear();
      return;
    }
    for (int a = 0; a < 3; a++) {
      glo[a] = lo[a];
      gdim[a] = hi[a] - lo[a] + 1;
    }
    egrid.assign((size_t)gdim[0] * gdim[1] * gdim[2], -1);
    for (Tile *t : dir.order)
      for (int l = 0; l < 64; l++) {
        int bx = t->c[0] * 4 + (l & 3), by = t->c[1] * 4 + ((l >> 2) & 3), bz = t->c[2] * 4 + (l >> 4);
        egrid[((size_t)(bz - glo[2]) * gdim[1] + (by - glo[1])) * gdim[0] + (bx - glo[0])] =
            entries[(size_t)t->slot * 64 + l];
      }
  }
  int entryAt(int bx, int by, int bz) const {
    int x = bx - glo[0], y = by - glo[1], z = bz - glo[2];
    if (x < 0 || y < 0 || z < 0 || x >= gdim[0] || y >= gdim[1] || z >= gdim[2]) return -1;
    return egrid[((size_t)z * gdim[1] + y) * gdim[0] + x];
  }
  int slotAtBrick(int bx, int by, int bz) const {
    int e = entryAt(bx, by, bz);
    return e < 0 ? -1 : pageSlot(e & 0xffffff);
  }

  // ---- cell access -------------------------------------------------------------------
  int maskF() const { return B - 1; }
  int maskV() const { return b - 1; }
  // Field cell (4 channels) of buffer `buf`, zero without a slot (fluid-sampling.wgsl:218-225).
  void fieldCell(const std::vector<R> &buf, int x, int y, int z, R out[4]) const {
    int s = slotAtBrick(x >> sF, y >> sF, z >> sF);
    if (s < 0) {
      out[0] = out[1] = out[2] = out[3] = 0;
      return;
    }
    const R *p = &buf[((size_t)s * nF + (((z & maskF()) * B + (y & maskF())) * B + (x & maskF()))) * 4];
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
    out[3] = p[3];
  }
  R fieldScalar(const std::vector<R> &buf, int x, int y, int z) const {
    int s = slotAtBrick(x >> sF, y >> sF, z >> sF);
    if (s < 0) return 0;
    return buf[(size_t)s * nF + (((z & maskF()) * B + (y & maskF())) * B + (x & maskF()))];
  }
  size_t vIndex(int s, int x, int y, int z) const {
    return (size_t)s * nV + (((z & maskV()) * b + (y & maskV())) * b + (x & maskV()));
  }
  R velCell(const std::vector<R> &buf, int c, int x, int y, int z) const {
    int s = slotAtBrick(x >> sV, y >> sV, z >> sV);
    if (s < 0) return 0;
    return buf[(size_t)s * nV * 3 + (size_t)c * nV + (((z & maskV()) * b + (y & maskV())) * b + (x & maskV()))];
  }
  R curlCell(int c, int x, int y, int z) const {
    int s = slotAtBrick(x >> sV, y >> sV, z >> sV);
    if (s < 0) return 0;
    return Curl[((size_t)s * nV + (((z & maskV()) * b + (y & maskV())) * b + (x & maskV()))) * 4 + c];
  }
  R velScalar(const std::vector<R> &buf, int x, int y, int z) const {
    int s = slotAtBrick(x >> sV, y >> sV, z >> sV);
    if (s < 0) return 0;
    return buf[vIndex(s, x, y, z)];
  }
  bool computes(int bx, int by, int bz) const { return entryAt(bx, by, bz) >= 0; }

  // Whether field cell `c` is solid (solidCell, solids.wgsl:15-32).
  bool solidCell(int x, int y, int z) const {
    if (P.ground && y < 0) return true;
    if (P.colliders == 0) return false;
    int s = slotAtBrick(x >> sF, y >> sF, z >> sF);
    if (s < 0) return false;
    int m = maskF();
    uint32_t bit = (uint32_t)((((z & m) << sF) | (y & m)) << sF | (x & m));
    size_t words = (size_t)nF / 32;
    return ((solids[(size_t)s * words + (bit >> 5)] >> (bit & 31u)) & 1u) != 0u;
  }
  // fieldCell + solidAt (solids.wgsl:9-36): the quarter-cell bias.
  bool solidAt(const R w[3]) const {
    int c[3];
    for (int a = 0; a < 3; a++) c[a] = (int)ffloor(w[a] * P.invHf + R(0.25));
    return solidCell(c[0], c[1], c[2]);
  }

  // Trilinear sample of the field grid at `pos` in cell units, centers at .5
  // (fluid-sampling.wgsl:235-264).
  void sampleField(const std::vector<R> &buf, const R pos[3], R out[4]) const {
    R q[3], f[3];
    int base[3];
    for (int a = 0; a < 3; a++) {
      q[a] = pos[a] - R(0.5);
      R fl = ffloor(q[a]);
      base[a] = (int)fl;
      f[a] = q[a] - fl;
    }
    out[0] = out[1] = out[2] = out[3] = 0;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          R wx = x ? f[0] : R(1) - f[0], wy = y ? f[1] : R(1) - f[1], wz = z ? f[2] : R(1) - f[2];
          R w = wx * wy * wz;
          R v[4];
          fieldCell(buf, base[0] + x, base[1] + y, base[2] + z, v);
          for (int c = 0; c < 4; c++) out[c] += v[c] * w;
        }
  }
  // The same for one velocity component.
  R sampleVel(const std::vector<R> &buf, int comp, const R pos[3]) const {
    R q[3], f[3];
    int base[3];
    for (int a = 0; a < 3; a++) {
      q[a] = pos[a] - R(0.5);
      R fl = ffloor(q[a]);
      base[a] = (int)fl;
      f[a] = q[a] - fl;
    }
    R out = 0;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          R wx = x ? f[0] : R(1) - f[0], wy = y ? f[1] : R(1) - f[1], wz = z ? f[2] : R(1) - f[2];
          out += velCell(buf, comp, base[0] + x, base[1] + y, base[2] + z) * (wx * wy * wz);
        }
    return out;
  }
  // velocityAt (fluid-advection.wgsl:10-21): each MAC component at its face offset.
  void velocityAt(const R w[3], R out[3]) const {
    for (int c = 0; c < 3; c++) {
      R pos[3];
      for (int a = 0; a < 3; a++) pos[a] = (a == c ? w[a] - R(0.5) * P.hv : w[a]) * P.invHv;
      out[c] = sampleVel(V, c, pos);
    }
  }
  // velocityOnFace (fluid-advection.wgsl:26-44).
  void velocityOnFace(const int id[3], int comp, const R orig[3], R out[3]) const {
    R fp[3];
    for (int a = 0; a < 3; a++) fp[a] = (R)id[a] + R(0.5);
    fp[comp] += R(0.5);
    for (int k = 0; k < 3; k++) {
      if (k == comp) {
        out[k] = orig[k];
        continue;
      }
      R pos[3] = {fp[0], fp[1], fp[2]};
      pos[k] = fp[k] - R(0.5);
      out[k] = sampleVel(V, k, pos);
    }
  }

  // ---- the step ----------------------------------------------------------------------
  void step(const fp_frame &frame, double dt) {
    if (!std::isfinite(dt) || dt <= 0 || dt > MAX_DT)
      throw Error("Simulation step must be in (0, 1/30]. Use fixed substeps.");
    substep(frame, dt);
    time += dt;
    steps++;
  }

  double tileSize() const { return TILE_BRICKS * B * voxel; }
  double velocitySpacing() const { return voxel * d; }

  // sourceBricks (FluidSimulation.ts:482-491).
  int sourceBricks(const fp_source &s, double dt) const {
    double size = B * voxel;
    double count = 1;
    for (int a = 0; a < 3; a++) {
      double motion = std::fabs(s.velocity[a]) + std::fabs(s.outward_speed);
      double pad = std::min(size, 2 * velocitySpacing() + motion * dt);
      count *= std::ceil((2 * (std::max(s.size[a], 0.03) + pad)) / size) + 1;
    }
    return (int)count;
  }

  void needBox(const double lo[3], const double hi[3], double margin) {
    double size = tileSize();
    int l[3], h[3];
    for (int a = 0; a < 3; a++) {
      l[a] = (int)std::floor((lo[a] - margin) / size);
      h[a] = (int)std::floor((hi[a] + margin) / size);
    }
    dir.need(l, h, steps);
  }

  // updateTiles (FluidSimulation.ts:873-912).
  void updateTiles(const fp_frame &f) {
    dir.retire(steps);
    double reach = 2 * B * voxel;
    double margin = reach + tileSize();
    std::vector<BrickBox> sourceBricks;
    for (int i = 0; i < f.source_count; i++) {
      const fp_source &s = f.sources[i];
      double lo[3], hi[3];
      if (s.shape == FP_SHAPE_MESH && s.samples && s.sample_count > 0) {
        // The bounds of a mesh source's samples (FluidSimulation.ts:880-894).
        for (int a = 0; a < 3; a++) {
          lo[a] = INFINITY;
          hi[a] = -INFINITY;
          for (int k = 0; k < s.sample_count; k++) {
            lo[a] = std::min(lo[a], s.samples[4 * (size_t)k + a]);
            hi[a] = std::max(hi[a], s.samples[4 * (size_t)k + a]);
          }
        }
        needBox(lo, hi, margin + voxel);
      } else {
        double extent = std::max(std::max(std::max(s.size[0], s.size[1]), s.size[2]), 0.03) +
                        s.outward_speed * 0.25;
        for (int a = 0; a < 3; a++) {
          lo[a] = s.position[a] - extent;
          hi[a] = s.position[a] + extent;
        }
        needBox(lo, hi, margin);
      }
      // bricksOver (FluidSimulation.ts:862-869)
      BrickBox bb;
      double size = B * voxel;
      for (int a = 0; a < 3; a++) {
        bb.lo[a] = (int)std::floor(lo[a] / size);
        bb.hi[a] = (int)std::floor(hi[a] / size);
      }
      sourceBricks.push_back(bb);
    }
    for (auto &ct : contentTiles) {
      int lo[3], hi[3];
      for (int a = 0; a < 3; a++) {
        lo[a] = ct.first[a] - 1;
        hi[a] = ct.first[a] + 1;
      }
      dir.need(lo, hi, steps);
    }
    if (dir.dirty) applyTiles();
    updateBoxes(sourceBricks);
  }

  static int floorDiv(int a, int b) { return (int)std::floor((double)a / b); }

  // updateBoxes (FluidSimulation.ts:913-969): per cluster of tiles, the union of the
  // bricks holding content at the latest measurement and of `extra`, grown by a margin
  // of two bricks inside the cluster's tiles; clusters in ascending order.
  void updateBoxes(const std::vector<BrickBox> &extra) {
    std::vector<int> order;
    std::vector<BrickBox> found;
    auto include = [&](const BrickBox &box) {
      int center[3];
      for (int a = 0; a < 3; a++) center[a] = floorDiv(box.lo[a] + box.hi[a], 2);
      if (ground) center[1] = std::max(0, center[1]);
      int slot = dir.slotAt(floorDiv(center[0], TILE_BRICKS), floorDiv(center[1], TILE_BRICKS),
                            floorDiv(center[2], TILE_BRICKS));
      if (slot < 0) return;
      int cluster = tables.tiles[(size_t)slot * 4 + 3];
      for (size_t i = 0; i < order.size(); i++)
        if (order[i] == cluster) {
          for (int a = 0; a < 3; a++) {
            found[i].lo[a] = std::min(found[i].lo[a], box.lo[a]);
            found[i].hi[a] = std::max(found[i].hi[a], box.hi[a]);
          }
          return;
        }
      order.push_back(cluster);
      found.push_back(box);
    };
    for (auto &ct : contentTiles) {
      BrickBox box;
      bool any = true;
      for (int a = 0; a < 3; a++) {
        uint32_t used = (ct.second >> (a * 4)) & 15u;
        if (!used) any = false;
        int lowest = 0, highest = 0;
        for (int k = 0; k < 4; k++)
          if (used >> k & 1u) {
            if (!(used & ((1u << k) - 1u))) lowest = k;
            highest = k;
          }
        box.lo[a] = ct.first[a] * TILE_BRICKS + lowest;
        box.hi[a] = ct.first[a] * TILE_BRICKS + highest;
      }
      if (any) include(box);
    }
    for (auto &b : extra) include(b);
    std::vector<int> clusters = order;
    std::sort(clusters.begin(), clusters.end());
    std::vector<BrickBox> boxes;
    for (int c : clusters) {
      size_t i = std::find(order.begin(), order.end(), c) - order.begin();
      const TileBox &tiles = tables.clusters[c];
      BrickBox out;
      for (int a = 0; a < 3; a++) {
        out.lo[a] = std::max(found[i].lo[a] - 2, tiles.lo[a] * TILE_BRICKS);
        out.hi[a] = std::min(found[i].hi[a] + 2, (tiles.hi[a] + 1) * TILE_BRICKS - 1);
      }
      boxes.push_back(out);
    }
    bool same = clusters == renderClusters && boxes.size() == renderBoxes.size();
    for (size_t i = 0; same && i < boxes.size(); i++)
      for (int a = 0; a < 3; a++)
        if (boxes[i].lo[a] != renderBoxes[i].lo[a] || boxes[i].hi[a] != renderBoxes[i].hi[a]) same = false;
    if (same && !renderStale) return;
    renderStale = false;
    renderBoxes = boxes;
    renderClusters = clusters;
    renderTiles = tables.tiles;
    for (size_t w = 3; w < renderTiles.size(); w += 4)
      if (renderTiles[w] >= 0) {
        auto it = std::find(clusters.begin(), clusters.end(), renderTiles[w]);
        renderTiles[w] = it == clusters.end() ? -1 : (int)(it - clusters.begin());
      }
    renderVersion++;
  }

  // growPool (FluidSimulation.ts:1013-1093).
  void growPool(int target) {
    int limit = slotLimit();
    int rows = poolRows(std::min(target, limit), side);
    int slots = std::min(limit, rows * side);
    if (slots <= poolSlots) return;
    int added = slots - poolSlots;
    resizePools(slots, true);
    std::vector<uint32_t> next(slots, 0u);
    for (int i = 0; i < poolSlots; i++) next[added + i] = freeSlots[i];
    for (int i = 0; i < added; i++) next[i] = (uint32_t)(poolSlots + added - 1 - i);
    freeSlots.swap(next);
    freeCount += added;
    poolSlots = slots;
    poolGrownAt = steps;
    geometryDirty = true;
  }

  // packFlameStep (flame.ts:43-87) with its validation.
  void packFlame(const fp_flame &fl, double dt) {
    auto finite = [](const char *name, double v) {
      float f = (float)v;
      if (!std::isfinite(v) || !std::isfinite(f)) throw Error(std::string(name) + " must be finite in f32.");
      return (double)f;
    };
    double stepf = finite("Flame timestep", dt), lifespan = finite("Flame lifespan", fl.lifespan);
    if (stepf <= 0 || lifespan <= 0) throw Error("Flame timestep and lifespan must be positive in f32.");
    finite("Lifetime decrement", stepf / lifespan);
    double smoke = finite("Smoke rate", fl.smoke_rate), heat = finite("Heat rate", fl.heat_rate);
    double expansion = finite("Expansion rate", fl.expansion_rate);
    if (smoke < 0 || heat < 0) throw Error("Smoke and heat rates must be nonnegative.");
    for (double rate : {smoke, heat, expansion}) finite("Integrated flame output", rate * lifespan);
    const char *names[3] = {"smoke", "heat", "expansion"};
    for (int ch = 0; ch < 3; ch++) {
      int n = fl.knot_count[ch];
      if (n < 2 || n > 8) throw Error(std::string(names[ch]) + " ramp needs 2–8 knots.");
      double previous = -1;
      for (int i = 0; i < n; i++) {
        double x = finite("lifetime", fl.knots[ch][i][0]), y = finite("response", fl.knots[ch][i][1]);
        if (x < 0 || x > 1 || y < 0 || y > 1 || x <= previous)
          throw Error(std::string(names[ch]) + " knots must increase in f32 with coordinates in [0,1].");
        if ((i == 0 && x != 0) || (i == n - 1 && x != 1))
          throw Error(std::string(names[ch]) + " ramp must span lifetime 0–1.");
        P.knots[ch][i][0] = (R)x;
        P.knots[ch][i][1] = (R)y;
        previous = x;
      }
      P.knotCount[ch] = n;
    }
    P.fdt = (R)stepf;
    P.lifespan = (R)lifespan;
    P.fsmoke = (R)smoke;
    P.fheat = (R)heat;
    P.fexpansion = (R)expansion;
  }

  // packScene (FluidSimulation.ts:1150-1366) for sphere/box emitters and colliders and
  // analytic forces. Returns whether the colliders' packed words changed.
  // MeshSourceRasterizer.prepare (MeshSourceRasterizer.ts:66-166): each sample spreads
  // weight * area over its eight trilinear field cells in units of the cells' mean face
  // area; a covered cell's mask min(1, coverage) scales the source's rates into it. Then
  // the brick tables (FluidSimulation.ts:1329-1351).
  void packMesh(const fp_frame &f) {
    using Cell = std::array<int, 3>;
    struct Acc {
      double flame = 0, heat = 0, smoke = 0, fuel = 0, motion[3] = {0, 0, 0}, response = 0;
    };
    std::map<Cell, Acc> cells;
    const double h = voxel, meanFace = (h * h + h * h + h * h) / 3;
    size_t visits = 0;
    for (int i = 0; i < f.source_count; i++) {
      const fp_source &s = f.sources[i];
      if (s.shape != FP_SHAPE_MESH) continue;
      if (!s.samples || s.sample_count <= 0 || !(s.surface_area > 0))
        throw Error("Mesh sources require surface samples and a positive surface area.");
      visits += (size_t)s.sample_count * 8;
      if (visits > 2000000) throw Error("Mesh source sampling exceeds two million cell visits.");
      std::map<Cell, double> coverage;
      for (int k = 0; k < s.sample_count; k++) {
        const double *p = s.samples + 4 * (size_t)k;
        double g[3], fr[3];
        int base[3];
        for (int a = 0; a < 3; a++) {
          g[a] = p[a] / h - 0.5;
          base[a] = (int)std::floor(g[a]);
          fr[a] = g[a] - base[a];
        }
        for (int z = 0; z < 2; z++)
          for (int y = 0; y < 2; y++)
            for (int x = 0; x < 2; x++) {
              double t = (x ? fr[0] : 1 - fr[0]) * (y ? fr[1] : 1 - fr[1]) * (z ? fr[2] : 1 - fr[2]);
              coverage[Cell{base[0] + x, base[1] + y, base[2] + z}] += p[3] * s.surface_area * t / meanFace;
            }
      }
      for (const auto &cv : coverage) {
        double mask = std::min(1.0, cv.second);
        if (mask <= 0) continue;
        Acc &c = cells[cv.first];
        c.flame = std::max(c.flame, s.flame * mask);
        c.heat += s.heat_rate * mask;
        c.smoke += s.smoke_rate * mask;
        c.fuel += s.fuel_rate * mask;
        double response = s.velocity_response * mask;
        for (int a = 0; a < 3; a++) c.motion[a] += s.velocity[a] * response;
        c.response += response;
      }
    }
    meshRecords.assign(8, R(0));
    meshCount = (int)cells.size();
    std::map<int, int> tables;
    std::vector<std::pair<size_t, int>> placed;
    int record = 1;
    for (const auto &c : cells) {
      const double v[8] = {c.second.flame, c.second.heat, c.second.smoke, c.second.fuel,
                           c.second.motion[0], c.second.motion[1], c.second.motion[2], c.second.response};
      for (double x : v) meshRecords.push_back((R)(float)x);
      const Cell &q = c.first;
      int id = brickIdOf(q[0] >> sF, q[1] >> sF, q[2] >> sF);
      if (id >= 0) {
        auto t = tables.emplace(id, (int)tables.size()).first->second;
        int m = B - 1;
        placed.push_back({(size_t)t * nF + (size_t)(((q[2] & m) * B + (q[1] & m)) * B + (q[0] & m)), record});
      }
      record++;
    }
    meshTables.assign((size_t)brickCount() + tables.size() * (size_t)nF, 0);
    std::fill(meshTables.begin(), meshTables.begin() + brickCount(), -1);
    for (const auto &t : tables) meshTables[t.first] = t.second;
    for (const auto &p : placed) meshTables[(size_t)brickCount() + p.first] = p.second;
  }
  // meshSourceAt (fluid-sources.wgsl:8-22): the record of the field cell at `w`
  // (fieldCell, solids.wgsl:9-11), 0 for none.
  int meshRecordAt(const R w[3]) const {
    int c[3];
    for (int a = 0; a < 3; a++) c[a] = (int)ffloor(w[a] * P.invHf + R(0.25));
    int id = brickIdOf(c[0] >> sF, c[1] >> sF, c[2] >> sF);
    if (id < 0 || meshTables[id] < 0) return 0;
    int m = B - 1;
    return meshTables[(size_t)brickCount() + (size_t)meshTables[id] * nF + ((c[2] & m) * B + (c[1] & m)) * B + (c[0] & m)];
  }

  bool packScene(const fp_frame &f) {
    packMesh(f);
    emitters.clear();
    for (int i = 0; i < f.source_count; i++) {
      const fp_source &s = f.sources[i];
      if (s.shape == FP_SHAPE_MESH) continue;
      Emitter<R> e;
      for (int a = 0; a < 3; a++) {
        e.pos[a] = (R)(float)s.position[a];
        e.size[a] = (R)(float)s.size[a];
        e.vel[a] = (R)(float)s.velocity[a];
      }
      e.box = s.shape == FP_SHAPE_BOX;
      e.flame = (R)(float)s.flame;
      e.heat = (R)(float)s.heat_rate;
      e.smoke = (R)(float)s.smoke_rate;
      e.response = (R)(float)s.velocity_response;
      e.outward = (R)(float)s.outward_speed;
      e.period = (R)(float)(s.has_variation ? s.variation_period : 0.5);
      e.strength = (R)(float)(s.has_variation ? s.variation_strength : 0.0);
      e.fuel = (R)(float)s.fuel_rate;
      for (int a = 0; a < 3; a++) e.invSize[a] = R(1) / fmax(e.size[a], R(0.03));
      e.invPeriod = R(1) / e.period;
      emitters.push_back(e);
    }
    // Colliders: 28 words each, compared with the last upload (BufferUploads.ts:6-16).
    int n = f.collider_count;
    std::vector<uint32_t> packed((size_t)std::max(1, n) * 28, 0u);
    std::vector<uint8_t> occupancy;
    colliders.clear();
    for (int i = 0; i < n; i++) {
      const fp_collider &c = f.colliders[i];
      if (c.shape == FP_SHAPE_MESH) {
        // A mesh collider's occupancy of its box of cells (FluidSimulation.ts:1221-1260).
        if (!c.occupancy || c.cells[0] <= 0 || c.cells[1] <= 0 || c.cells[2] <= 0)
          throw Error("Mesh collider occupancy does not match its size.");
        Collider<R> cr{};
        cr.box = 2;
        size_t cells = (size_t)c.cells[0] * c.cells[1] * c.cells[2];
        cr.inside.resize(cells);
        for (size_t k = 0; k < cells; k++) cr.inside[k] = c.occupancy[k] != 0;
        for (int a = 0; a < 3; a++) {
          cr.lo[a] = c.lo[a];
          cr.size[a] = c.cells[a];
          packed[(size_t)i * 28 + 20 + a] = (uint32_t)c.lo[a];
          packed[(size_t)i * 28 + 24 + a] = (uint32_t)c.cells[a];
        }
        packed[(size_t)i * 28 + 3] = 2;
        occupancy.insert(occupancy.end(), cr.inside.begin(), cr.inside.end());
        colliders.push_back(cr);
        continue;
      }
      float w[20] = {0};
      for (int a = 0; a < 3; a++) {
        w[a] = (float)c.position[a];
        w[4 + a] = (float)c.size[a];
        for (int k = 0; k < 3; k++) w[8 + 4 * k + a] = (float)c.axes[k][a];
      }
      w[3] = c.shape == FP_SHAPE_BOX ? 1.0f : 0.0f;
      std::memcpy(&packed[(size_t)i * 28], w, sizeof(w));
      Collider<R> cr{};
      for (int a = 0; a < 3; a++) {
        cr.center[a] = (R)w[a];
        cr.extent[a] = (R)w[4 + a];
        for (int k = 0; k < 3; k++) cr.axes[k][a] = (R)w[8 + 4 * k + a];
      }
      cr.box = c.shape == FP_SHAPE_BOX;
      colliders.push_back(cr);
    }
    bool changed = packed != prevColliders || n != colliderCount || occupancy != prevOccupancy;
    prevColliders = packed;
    prevOccupancy = occupancy;
    colliderCount = n;
    // Forces (ForceFields.ts:41-114), analytic targets only.
    forces.clear();
    for (int i = 0; i < f.force_count; i++) {
      const fp_force &fc = f.forces[i];
      Force<R> F_;
      for (int a = 0; a < 3; a++) {
        F_.axis[a] = (R)(float)fc.vector[a];
        F_.center[a] = (R)(float)fc.center[a];
      }
      F_.strength = (R)(float)fc.strength;
      F_.kind = (R)fc.type;
      F_.scale = (R)(float)fc.scale;
      F_.lift = (R)(float)fc.lift;
      F_.inward = (R)(float)fc.inward;
      F_.targeted = fc.targeted != 0;
      for (int k = 0; k < fc.source_count; k++) {
        const fp_source &s = fc.sources[k];
        if (s.shape == FP_SHAPE_MESH) {
          // The velocity cells the target's samples cover (ForceFields.ts:114-179): coverage
          // above zero on the velocity lattice, in bricks that have ids this step.
          if (!s.samples || s.sample_count <= 0 || !(s.surface_area > 0))
            throw Error("Mesh sources require surface samples and a positive surface area.");
          F_.meshTargets = true;
          const double h = velocitySpacing(), meanFace = (h * h + h * h + h * h) / 3;
          std::map<std::array<int, 3>, double> cov;
          for (int q = 0; q < s.sample_count; q++) {
            const double *p = s.samples + 4 * (size_t)q;
            double fr[3];
            int base[3];
            for (int a = 0; a < 3; a++) {
              double g = p[a] / h - 0.5;
              base[a] = (int)std::floor(g);
              fr[a] = g - base[a];
            }
            for (int z = 0; z < 2; z++)
              for (int y = 0; y < 2; y++)
                for (int x = 0; x < 2; x++) {
                  double t = (x ? fr[0] : 1 - fr[0]) * (y ? fr[1] : 1 - fr[1]) * (z ? fr[2] : 1 - fr[2]);
                  cov[{base[0] + x, base[1] + y, base[2] + z}] += p[3] * s.surface_area * t / meanFace;
                }
          }
          for (const auto &c : cov)
            if (std::min(1.0, c.second) > 0 && brickIdOf(c.first[0] >> sV, c.first[1] >> sV, c.first[2] >> sV) >= 0)
              F_.covered[c.first] = true;
          continue;
        }
        Region<R> r;
        for (int a = 0; a < 3; a++) {
          r.center[a] = (R)(float)s.position[a];
          r.size[a] = (R)(float)s.size[a];
        }
        r.kind = s.shape == FP_SHAPE_BOX ? R(3) : R(1);
        for (int a = 0; a < 3; a++) r.invSize[a] = R(1) / fmax(r.size[a], R(0.03));
        F_.regions.push_back(r);
      }
      forces.push_back(F_);
    }
    return changed;
  }

  void setParams(const fp_frame &f, double dt) {
    const fp_settings &s = f.simulation;
    P.hf = (R)(float)voxel;
    P.hv = (R)(float)velocitySpacing();
    P.invHf = R(1) / P.hf;
    P.invHv = R(1) / P.hv;
    P.hs = (R)(float)(voxel * smokeDiv);
    P.invHs = R(1) / P.hs;
    P.dt = (R)(float)dt;
    P.buoyancy = (R)(float)s.buoyancy;
    P.cooling = (R)(float)s.cooling;
    P.smokeWeight = (R)(float)s.smoke_weight;
    P.damping = (R)(float)s.velocity_damping;
    P.dissipation = (R)(float)s.dissipation;
    P.vorticity = (R)(float)s.vorticity;
    P.cutoff = (R)(float)cutoff;
    P.lifetimeZero = !mayHaveFlame;
    P.time = (R)(float)time;
    float seedf = (float)s.seed;
    P.seed = (R)seedf;
    std::memcpy(&P.seedBits, &seedf, 4);
    P.colliders = colliderCount;
    P.ground = ground;
    fp_fuel fuel = f.has_fuel ? f.fuel : fp_fuel{0, 0.5, 4.0};
    P.ignition = (R)(float)fuel.ignition_heat;
    P.burnRate = (R)(float)fuel.burn_rate;
    P.fuelEnabled = fuel.enabled != 0;
  }

  void substep(const fp_frame &f, double dt) {
    if (dir.ground != ground) {
      dir.ground = ground;
      geometryDirty = true;
    }
    updateTiles(f);
    int wanted = 0;
    for (int i = 0; i < f.source_count; i++)
      if (std::find(sourceIds.begin(), sourceIds.end(), f.sources[i].id) == sourceIds.end())
        wanted += sourceBricks(f.sources[i], dt);
    sourceIds.clear();
    for (int i = 0; i < f.source_count; i++) sourceIds.push_back(f.sources[i].id);
    if (wanted > 0)
      poolTarget = std::max(poolTarget, (int)std::ceil((usedSlots + wanted) * 1.5));
    if (poolTarget > poolSlots) growPool(poolTarget);
    bool solidsChanged = packScene(f);
    if (solidsChanged) {
      solidsStale = true;
      geometryDirty = true;
    }
    packFlame(f.combustion, dt);
    fp_fuel fuel = f.has_fuel ? f.fuel : fp_fuel{0, 0.5, 4.0};
    bool anyFlame = false;
    for (int i = 0; i < f.source_count; i++) anyFlame = anyFlame || f.sources[i].flame > 0;
    mayHaveFlame = mayHaveFlame || fuel.enabled || anyFlame;
    setParams(f, dt);
    for (int t = 0; t < capacity; t++) activity[(size_t)brickCount() + t] = 0;
    allocation(f.collider_count > 0);
    buildEntryGrid();
    zeroFresh();
    advectVelocity();
    correctVelocity();
    std::vector<uint32_t> support;
    advectScalars();
    evolveFields();
    std::swap(F, Fn);
    if (coarse()) {
      advectSmoke();
      evolveSmoke();
      std::swap(S, Sn);
    }
    if (P.vorticity != R(0)) computeCurl();
    applyForces();
    computeDivergence();
    solvePressure();
    project();
    readActivity();
  }

  // ---- allocation (fluid-allocation.wgsl:129-387) ------------------------------------
  std::vector<uint32_t> st;
  bool sourceNear(const int bk[3]) const {
    R size = (R)(float)(1 << sF) * P.hf;
    R lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      lo[a] = (R)bk[a] * size;
      hi[a] = ((R)bk[a] + R(1)) * size;
    }
    for (const Emitter<R> &e : emitters) {
      bool in = true;
      for (int a = 0; a < 3; a++) {
        R motion = fabs_(e.vel[a]) + fabs_(e.outward);
        R pad = fmin(size, R(2) * P.hv + motion * P.dt);
        R extent = fmax(e.size[a], R(0.03)) + pad;
        if (!(e.pos[a] + extent > lo[a] && e.pos[a] - extent < hi[a])) in = false;
      }
      if (in) return true;
    }
    return false;
  }

  int heldSlots() const { return std::max(0, poolSlots - slotLimit()); }

  void allocation(bool haveColliders) {
    int n = brickCount();
    st.assign(n, 0u);
    activeList.clear();
    zeroList.clear();
    // buildBricks (:129-175)
    for (int index = 0; index < n; index++) {
      uint32_t state = 0;
      int ts = tileState(index >> 6);
      if (ts != TILE_FREE) {
        int bk[3];
        brickOf(index, bk);
        bool open = ts != TILE_CLOSING;
        bool listed = false;
        for (int z = -1; z <= 1 && open && !listed; z++)
          for (int y = -1; y <= 1 && !listed; y++)
            for (int x = -1; x <= 1 && !listed; x++) {
              int nb = brickIdOf(bk[0] + x, bk[1] + y, bk[2] + z);
              if (nb < 0) continue;
              int toward = (-z + 1) * 9 + (-y + 1) * 3 + (-x + 1);
              listed = ((activity[nb] >> 3) & (1u << toward)) != 0u || (meshCount > 0 && meshTables[nb] >= 0);
            }
        if (open && !listed) listed = sourceNear(bk);
        int page = pages[index];
        bool was = (brickState[index] & 1u) != 0u;
        state = (uint32_t)listed | ((uint32_t)was << 2);
        if (!listed && page >= 0) state |= BRICK_FREES;
        if (listed && page < 0) state |= BRICK_REQUESTS;
      }
      st[index] = state;
    }
    // freeBricks (:179-195): pushed in brick order above the stack.
    int freed = 0;
    for (int index = 0; index < n; index++) {
      if (!(st[index] & BRICK_FREES)) continue;
      int page = pages[index];
      freeSlots[(size_t)freeCount + freed] = (uint32_t)pageSlot(page);
      freed++;
      if (st[index] & 4u) zeroList.push_back({index, (uint32_t)page});
      pages[index] = -1;
    }
    // allocateBricks (:200-230)
    int held = heldSlots();
    int requests = 0;
    for (int index = 0; index < n; index++) {
      uint32_t state = st[index];
      int page = pages[index];
      bool fresh = false;
      if (state & BRICK_REQUESTS) {
        int top = freeCount + freed - requests;
        requests++;
        if (top > held) {
          page = slotPage((int)freeSlots[top - 1]);
          pages[index] = page;
          zeroList.push_back({index, (uint32_t)page | FRESH_SLOT});
          fresh = true;
        }
      }
      if (page >= 0) activeList.push_back(index);
      brickState[index] = page >= 0 ? 1u : 0u;
      entries[index] = page >= 0 ? (page | (fresh ? FRESH_ENTRY : 0)) : -1;
      uint32_t flags = activity[index];
      activity[index] = (flags & EXPANDS) ? MAY_EXPAND : 0u;
    }
    // linkBricks (:234-268)
    for (int index = 0; index < n; index++) {
      if (entries[index] < 0) continue;
      int bk[3];
      brickOf(index, bk);
      int32_t *w = &records[(size_t)index * BRICK_RECORD_WORDS];
      for (int k = 0; k < 27; k++) {
        int nb = brickIdOf(bk[0] + k % 3 - 1, bk[1] + (k / 3) % 3 - 1, bk[2] + k / 9 - 1);
        w[k] = nb >= 0 ? entries[nb] : -1;
      }
      int unknowns = 1;
      for (int k = 0; k < 8; k++) {
        int dx = -(k & 1), dy = -((k >> 1) & 1), dz = -((k >> 2) & 1);
        if (ground && bk[1] + dy < 0) continue;
        if (w[(dz + 1) * 9 + (dy + 1) * 3 + dx + 1] < 0) unknowns = 0;
      }
      w[PRESSURE_FLAG] = unknowns;
    }
    // finishBricks (:352-387)
    int freeAfter = freeCount + freed;
    int granted = std::min(requests, std::max(freeAfter - held, 0));
    freeCount = freeAfter - granted;
    failures += (uint32_t)(requests - granted);
    // buildSolids (:292-341) for fresh slots, or every slot after the colliders changed.
    if (haveColliders || solidsStale) buildSolids();
    solidsStale = false;
  }

  bool insideCollider(const Collider<R> &c, int x, int y, int z) const {
    if (c.box == 2) {
      int q[3] = {x - c.lo[0], y - c.lo[1], z - c.lo[2]};
      for (int a = 0; a < 3; a++)
        if (q[a] < 0 || q[a] >= c.size[a]) return false;
      return c.inside[((size_t)q[2] * c.size[1] + q[1]) * c.size[0] + q[0]] != 0;
    }
    R w[3] = {((R)x + R(0.5)) * P.hf, ((R)y + R(0.5)) * P.hf, ((R)z + R(0.5)) * P.hf};
    R rel[3] = {w[0] - c.center[0], w[1] - c.center[1], w[2] - c.center[2]};
    R dd[3];
    for (int k = 0; k < 3; k++) {
      R proj = c.axes[k][0] * rel[0] + c.axes[k][1] * rel[1] + c.axes[k][2] * rel[2];
      dd[k] = fabs_(proj / c.extent[k]);
    }
    if (c.box == 1) return dd[0] < R(1) && dd[1] < R(1) && dd[2] < R(1);
    return dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] < R(1);
  }
  void drawSolids(int id, int page) {
    int bk[3];
    brickOf(id, bk);
    int slot = pageSlot(page);
    size_t words = (size_t)nF / 32;
    for (size_t word = 0; word < words; word++) {
      uint32_t bits = 0;
      for (int k = 0; k < 32; k++) {
        uint32_t bit = (uint32_t)word * 32 + k;
        int x = bk[0] * B + (int)(bit & (B - 1)), y = bk[1] * B + (int)((bit >> sF) & (B - 1)),
            z = bk[2] * B + (int)(bit >> (2 * sF));
        for (const auto &c : colliders)
          if (insideCollider(c, x, y, z)) {
            bits |= 1u << k;
            break;
          }
      }
      solids[(size_t)slot * words + word] = bits;
    }
  }
  void buildSolids() {
    if (solidsStale) {
      for (int id = 0; id < brickCount(); id++) {
        int e = entries[id];
        if (e >= 0) drawSolids(id, e & 0xffffff);
      }
      return;
    }
    for (auto &z : zeroList)
      if (z.second & FRESH_SLOT) drawSolids(z.first, (int)(z.second & 0xffffffu));
  }

  // zeroVelocityBricks / zeroFieldBricks (fluid-zeroing.wgsl:76-105): a fresh slot is
  // still, empty air. A freed slot not granted again holds no brick, so nothing reads it.
  void zeroFresh() {
    for (auto &z : zeroList) {
      if (!(z.second & FRESH_SLOT)) continue;
      int s = pageSlot((int)(z.second & 0xffffffu));
      std::fill(F.begin() + (size_t)s * nF * 4, F.begin() + (size_t)(s + 1) * nF * 4, R(0));
      std::fill(E.begin() + (size_t)s * nF, E.begin() + (size_t)(s + 1) * nF, R(0));
      std::fill(V.begin() + (size_t)s * nV * 3, V.begin() + (size_t)(s + 1) * nV * 3, R(0));
      // zeroSmokeBricks (smoke.wgsl:187-223)
      if (coarse()) std::fill(S.begin() + (size_t)s * nS, S.begin() + (size_t)(s + 1) * nS, R(0));
    }
  }

  // The per-cell passes below run their bricks on the host's threads (OpenMP): a pass
  // writes only its own cells and its own brick's flags, so results do not depend on the
  // thread count.
  template <class Fn> void forActiveVelocityCells(Fn fn) {
    const long count = (long)activeList.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < count; i++) {
      const int id = activeList[(size_t)i];
      int bk[3];
      brickOf(id, bk);
      int s = pageSlot(entries[id] & 0xffffff);
      for (int l = 0; l < nV; l++) {
        int c[3] = {bk[0] * b + (l & (b - 1)), bk[1] * b + ((l >> sV) & (b - 1)), bk[2] * b + (l >> (2 * sV))};
        fn(s, l, c);
      }
    }
  }
  template <class Fn> void forActiveFieldCells(Fn fn) {
    const long count = (long)activeList.size();
#pragma omp parallel for schedule(dynamic, 2)
    for (long i = 0; i < count; i++) {
      const int id = activeList[(size_t)i];
      int bk[3];
      brickOf(id, bk);
      int s = pageSlot(entries[id] & 0xffffff);
      for (int l = 0; l < nF; l++) {
        int c[3] = {bk[0] * B + (l & (B - 1)), bk[1] * B + ((l >> sF) & (B - 1)), bk[2] * B + (l >> (2 * sF))};
        fn(id, s, l, c);
      }
    }
  }

  // ---- velocity transport (fluid-velocity-transport.wgsl) ---------------------------
  void centerV(const int id[3], R w[3]) const {
    for (int a = 0; a < 3; a++) w[a] = ((R)id[a] + R(0.5)) * P.hv;
  }
  // advectVelocityComponent (fluid-velocity-transport.wgsl:28-41) → value and departure base cell.
  R advectComponent(const int id[3], const R pos[3], int c, const R orig[3], int donor[3]) const {
    R off[3] = {0, 0, 0};
    off[c] = R(0.5) * P.hv;
    R fp[3], fv[3], mid[3], dep[3], v[3];
    for (int a = 0; a < 3; a++) fp[a] = pos[a] + off[a];
    velocityOnFace(id, c, orig, fv);
    for (int a = 0; a < 3; a++) mid[a] = fp[a] - R(0.5) * P.dt * fv[a];
    velocityAt(mid, v);
    for (int a = 0; a < 3; a++) dep[a] = fp[a] - P.dt * v[a];
    R sp[3];
    for (int a = 0; a < 3; a++) {
      R rel = dep[a] - off[a];
      donor[a] = (int)ffloor(rel * P.invHv - R(0.5));
      sp[a] = rel * P.invHv;
    }
    return sampleVel(V, c, sp);
  }
  void advectVelocity() {
    forActiveVelocityCells([&](int s, int l, const int c[3]) {
      R pos[3];
      centerV(c, pos);
      R *out = &Vadv[(size_t)s * nV * 3];
      if (solidAt(pos)) {
        out[l] = out[nV + l] = out[2 * nV + l] = 0;
        return;
      }
      R orig[3] = {V[(size_t)s * nV * 3 + l], V[(size_t)s * nV * 3 + nV + l], V[(size_t)s * nV * 3 + 2 * nV + l]};
      int donor[3];
      for (int k = 0; k < 3; k++) out[k * nV + l] = advectComponent(c, pos, k, orig, donor);
    });
  }
  // Donor bounds (:165-200, :204-212): the range of a component over the 2^3 cells from
  // `base`, or [0, 0] when base's brick has no slot.
  void boundsAt(const int base[3], int comp, R &lo, R &hi) const {
    if (slotAtBrick(base[0] >> sV, base[1] >> sV, base[2] >> sV) < 0) {
      lo = hi = 0;
      return;
    }
    lo = R(1e5);
    hi = R(-1e5);
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          R v = velCell(V, comp, base[0] + x, base[1] + y, base[2] + z);
          lo = fmin(lo, v);
          hi = fmax(hi, v);
        }
  }
  // correctVelocityComponent (fluid-velocity-transport.wgsl:114-155).
  R correctComponent(const int id[3], const R pos[3], int c, const R orig[3], R predicted) const {
    R off[3] = {0, 0, 0};
    off[c] = R(0.5) * P.hv;
    R fp[3];
    for (int a = 0; a < 3; a++) fp[a] = pos[a] + off[a];
    // The departure base, as advection found it.
    R fv[3], mid[3], v[3];
    velocityOnFace(id, c, orig, fv);
    for (int a = 0; a < 3; a++) mid[a] = fp[a] - R(0.5) * P.dt * fv[a];
    velocityAt(mid, v);
    int base[3];
    for (int a = 0; a < 3; a++) base[a] = (int)ffloor((fp[a] - P.dt * v[a] - off[a]) * P.invHv - R(0.5));
    R lo, hi;
    boundsAt(base, c, lo, hi);
    if (lo == hi && predicted == lo) return predicted;
    R fmid[3], fw[3], sp[3];
    for (int a = 0; a < 3; a++) fmid[a] = fp[a] + R(0.5) * P.dt * fv[a];
    velocityAt(fmid, fw);
    for (int a = 0; a < 3; a++) sp[a] = (fp[a] + P.dt * fw[a] - off[a]) * P.invHv;
    R reversed = sampleVel(Vadv, c, sp);
    R corrected = predicted + R(0.5) * (orig[c] - reversed);
    return (corrected < lo || corrected > hi) ? predicted : corrected;
  }
  void correctVelocity() {
    forActiveVelocityCells([&](int s, int l, const int c[3]) {
      R pos[3];
      centerV(c, pos);
      R *out = &Vcor[(size_t)s * nV * 3];
      if (solidAt(pos)) {
        out[l] = out[nV + l] = out[2 * nV + l] = 0;
        return;
      }
      const R *vs = &V[(size_t)s * nV * 3];
      const R *as = &Vadv[(size_t)s * nV * 3];
      R orig[3] = {vs[l], vs[nV + l], vs[2 * nV + l]};
      for (int k = 0; k < 3; k++) out[k * nV + l] = correctComponent(c, pos, k, orig, as[k * nV + l]);
    });
  }

  // ---- scalar transport and evolution (fluid-scalar-transport.wgsl,
  //      fluid-field-evolution.wgsl) -----------------------------------------------------
  void centerF(const int id[3], R w[3]) const {
    for (int a = 0; a < 3; a++) w[a] = ((R)id[a] + R(0.5)) * P.hf;
  }
  void traceCenter(const R pos[3], R cv[3]) const { velocityAt(pos, cv); }
  void traceBack(const R pos[3], const R cv[3], R out[3]) const {
    R m[3], v[3];
    for (int a = 0; a < 3; a++) m[a] = pos[a] - R(0.5) * P.dt * cv[a];
    velocityAt(m, v);
    for (int a = 0; a < 3; a++) out[a] = pos[a] - P.dt * v[a];
  }
  void traceForward(const R pos[3], const R cv[3], R out[3]) const {
    R m[3], v[3];
    for (int a = 0; a < 3; a++) m[a] = pos[a] + R(0.5) * P.dt * cv[a];
    velocityAt(m, v);
    for (int a = 0; a < 3; a++) out[a] = pos[a] + P.dt * v[a];
  }
  struct Donors {
    R mn[3], mx[3], lmn, lmx;  // smoke, heat, fuel; lifetime
  };
  // predictScalarsWithBounds (:38-62).
  void predict(const int base[3], const R frac[3], R value[4], Donors &dn) const {
    for (int k = 0; k < 3; k++) {
      dn.mn[k] = R(65504);
      dn.mx[k] = R(-65504);
    }
    dn.lmn = 1;
    dn.lmx = 0;
    value[0] = value[1] = value[2] = value[3] = 0;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          R smp[4];
          fieldCell(F, base[0] + x, base[1] + y, base[2] + z, smp);
          const int ch[3] = {0, 1, 3};
          for (int k = 0; k < 3; k++) {
            dn.mn[k] = fmin(dn.mn[k], smp[ch[k]]);
            dn.mx[k] = fmax(dn.mx[k], smp[ch[k]]);
          }
          dn.lmn = fmin(dn.lmn, smp[2]);
          dn.lmx = fmax(dn.lmx, smp[2]);
          R wx = x ? frac[0] : R(1) - frac[0], wy = y ? frac[1] : R(1) - frac[1], wz = z ? frac[2] : R(1) - frac[2];
          for (int c = 0; c < 4; c++) value[c] += smp[c] * wx * wy * wz;
        }
  }
  // The departure cell of a backward trace in cell coordinates (:134-136).
  void departureCell(const int id[3], const R pos[3], const R dep[3], int base[3], R frac[3]) const {
    for (int a = 0; a < 3; a++) {
      R cell = (R)id[a] + (dep[a] - pos[a]) * P.invHf;
      R fl = ffloor(cell);
      base[a] = (int)fl;
      frac[a] = cell - fl;
    }
  }
  void advectScalars() {
    forActiveFieldCells([&](int, int s, int l, const int c[3]) {
      R pos[3], cv[3], dep[3];
      centerF(c, pos);
      traceCenter(pos, cv);
      traceBack(pos, cv, dep);
      R *out = &Fp[((size_t)s * nF + l) * 4];
      if (!maccormack) {
        R sp[3], v[4];
        for (int a = 0; a < 3; a++) sp[a] = dep[a] * P.invHf;
        sampleField(F, sp, v);
        out[0] = coarse() ? R(0) : v[0];
        out[1] = v[1];
        out[2] = clampv(v[2], R(0), R(1));
        out[3] = v[3];
        return;
      }
      int base[3];
      R frac[3], v[4];
      departureCell(c, pos, dep, base, frac);
      Donors dn;
      predict(base, frac, v, dn);
      // Coarse smoke: the fine X channel only stages the step's smoke production
      // (fluid-scalar-transport.wgsl:127, :137, :191-199).
      out[0] = coarse() ? R(0) : v[0];
      out[1] = v[1];
      out[2] = P.lifetimeZero ? R(0) : v[2];
      out[3] = v[3];
    });
  }
  // correctedFineScalars (:144-188); bounds are recomputed from the same backward trace
  // that advection used, which is what the reference caches in its donor words.
  void correctedScalars(const int id[3], int s, int l, const R pos[3], const R cv[3], R out[4]) const {
    const R *adv = &Fp[((size_t)s * nF + l) * 4];
    if (!maccormack) {
      for (int k = 0; k < 4; k++) out[k] = adv[k];
      return;
    }
    R dep[3];
    traceBack(pos, cv, dep);
    int base[3];
    R frac[3], tmp[4];
    departureCell(id, pos, dep, base, frac);
    Donors dn;
    predict(base, frac, tmp, dn);
    bool same = dn.mn[0] == dn.mx[0] && dn.mn[1] == dn.mx[1] && dn.mn[2] == dn.mx[2] && dn.lmn == dn.lmx;
    if (same) {
      out[0] = dn.mn[0];
      out[1] = dn.mn[1];
      out[2] = clampv(dn.lmn, R(0), R(1));
      out[3] = dn.mn[2];
      return;
    }
    R fw[3];
    traceForward(pos, cv, fw);
    const R *orig = &F[((size_t)s * nF + l) * 4];
    R lifetime = adv[2];
    R rev[3];
    uint8_t decided = 16;
    if (dn.lmn != dn.lmx || adv[2] != dn.lmn) {
      // sampleScalarsExplicit (:6-21) at the forward cell.
      R cell[3], fr[3];
      int bs[3];
      for (int a = 0; a < 3; a++) {
        cell[a] = (R)id[a] + (fw[a] - pos[a]) * P.invHf;
        R fl = ffloor(cell[a]);
        bs[a] = (int)fl;
        fr[a] = cell[a] - fl;
      }
      R val[4] = {0, 0, 0, 0};
      for (int z = 0; z < 2; z++)
        for (int y = 0; y < 2; y++)
          for (int x = 0; x < 2; x++) {
            R smp[4];
            fieldCell(Fp, bs[0] + x, bs[1] + y, bs[2] + z, smp);
            R wx = x ? fr[0] : R(1) - fr[0], wy = y ? fr[1] : R(1) - fr[1], wz = z ? fr[2] : R(1) - fr[2];
            for (int k = 0; k < 4; k++) val[k] += smp[k] * wx * wy * wz;
          }
      rev[0] = val[0];
      rev[1] = val[1];
      rev[2] = val[3];
      R cl = adv[2] + R(0.5) * (orig[2] - val[2]);
      lifetime = (cl < dn.lmn || cl > dn.lmx) ? adv[2] : cl;
      decided |= (cl < dn.lmn || cl > dn.lmx) ? 32 : 32 | 4;
    } else {
      R sp[3], val[4];
      for (int a = 0; a < 3; a++) sp[a] = fw[a] * P.invHf;
      sampleField(Fp, sp, val);
      rev[0] = val[0];
      rev[1] = val[1];
      rev[2] = val[3];
    }
    const int ch[3] = {0, 1, 3};
    R bounded[3];
    for (int k = 0; k < 3; k++) {
      R corr = adv[ch[k]] + R(0.5) * (orig[ch[k]] - rev[k]);
      bounded[k] = (corr < dn.mn[k] || corr > dn.mx[k]) ? adv[ch[k]] : corr;
      if (!(corr < dn.mn[k] || corr > dn.mx[k])) decided |= (uint8_t)(1 << ch[k]);
    }
    limiter[(size_t)s * nF + l] = decided;
    out[0] = bounded[0];
    out[1] = bounded[1];
    out[2] = clampv(lifetime, R(0), R(1));
    out[3] = bounded[2];
  }

  // contentSupport (fluid-content.wgsl:29-39) on the field grid.
  uint32_t contentSupport(const int id[3], const R vel[3]) const {
    return supportOn(B, P.hv / P.hf, P.invHf, id, vel);
  }
  // The same on a grid of bricks `side` cells a side, velocity cells `ratio` of its cells
  // wide and cells 1 / invH wide.
  uint32_t supportOn(int side, R ratio, R invH, const int id[3], const R vel[3]) const {
    R sidef = (R)side;
    bool low[3], high[3];
    for (int a = 0; a < 3; a++) {
      R local = (R)(id[a] & (side - 1)) + R(0.5);
      R pad = fmin(sidef, fmax(R(2), R(2) * ratio) + fabs_(vel[a]) * P.dt * invH);
      low[a] = local < pad;
      high[a] = local + pad >= sidef;
    }
    uint32_t row = 2u | (low[0] ? 1u : 0u) | (high[0] ? 4u : 0u);
    uint32_t plane = (row << 3) | (low[1] ? row : 0u) | (high[1] ? row << 6 : 0u);
    return (plane << 9) | (low[2] ? plane : 0u) | (high[2] ? plane << 18 : 0u);
  }

  void evolveFields() {
    std::vector<uint32_t> busy(brickCount(), 0u);
    std::vector<uint8_t> expands(brickCount(), 0);
    forActiveFieldCells([&](int bid, int s, int l, const int c[3]) {
      R pos[3];
      centerF(c, pos);
      R value[4] = {0, 0, 0, 0};
      R expansion = 0;
      bool injected = false;
      R cv[3];
      traceCenter(pos, cv);
      limiter[(size_t)s * nF + l] = 0;
      if (!solidAt(pos)) {
        R tr[4];
        correctedScalars(c, s, l, pos, cv, tr);
        if (coarse()) tr[0] = 0;
        R lifetime = tr[2];
        R sp[2] = {0, 0};
        R injectedFuel = 0;
        for (const Emitter<R> &e : emitters) {
          R wgt = emitterWeight(e, pos);
          if (wgt == R(0)) continue;
          R var[2];
          burstVariation(P.seedBits, e, pos, var);
          lifetime = fmax(lifetime, e.flame * wgt * var[1]);
          sp[0] += e.smoke * wgt * P.dt;
          sp[1] += e.heat * var[0] * wgt * P.dt;
          injectedFuel += e.fuel * wgt * P.dt;
        }
        if (meshCount > 0) {
          const R *mesh = &meshRecords[8 * (size_t)meshRecordAt(pos)];
          lifetime = fmax(lifetime, mesh[0]);
          sp[0] += mesh[2] * P.dt;
          sp[1] += mesh[1] * P.dt;
          injectedFuel += mesh[3] * P.dt;
        }
        R remaining = 0, burned = 0;
        if (P.fuelEnabled) {
          remaining = tr[3] + injectedFuel;
          if (tr[1] + sp[1] >= P.ignition) {
            lifetime = fmax(lifetime, fmin(remaining, R(1)));
            R unburned = remaining * fexp(-P.burnRate * P.dt);
            burned = remaining - unburned;
            remaining = unburned;
          }
        }
        R after, prod[3];
        evaluateFlame(P, clampv(lifetime, R(0), R(1)), after, prod);
        expansion = prod[2] + P.fexpansion * burned / P.dt;
        R lossS = fexp(-P.dissipation * P.dt), lossH = fexp(-P.cooling * P.dt);
        R halfS = fexp(-P.dissipation * P.dt * R(0.5)), halfH = fexp(-P.cooling * P.dt * R(0.5));
        R dens = tr[0] * lossS + (sp[0] + prod[0]) * halfS;
        R heat = tr[1] * lossH + (sp[1] + prod[1]) * halfH;
        value[0] = fmax(dens, R(0));
        value[1] = fmax(heat, R(0));
        value[2] = after;
        value[3] = remaining;
        injected = sp[0] > R(0) || sp[1] > R(0) || lifetime > tr[2] || injectedFuel > R(0);
      }
      // clearDormant (fluid-common.wgsl:8-13)
      if (!injected && value[2] == R(0) && value[0] < R(0.001) && value[1] < R(0.001) && value[3] < R(0.001))
        value[0] = value[1] = value[2] = value[3] = 0;
      R *o = &Fn[((size_t)s * nF + l) * 4];
      for (int k = 0; k < 4; k++) o[k] = value[k];
      E[(size_t)s * nF + l] = expansion;
      if (expansion != R(0)) expands[bid] = 1;
      // holdsContent (fluid-content.wgsl:22-24)
      if (fmax(fmax(coarse() ? R(0) : value[0], value[2]), value[3]) > P.cutoff) busy[bid] |= contentSupport(c, cv);
    });
    // recordContent (fluid-content.wgsl:43-64)
    for (int id : activeList) {
      uint32_t flags = 0;
      if (busy[id]) {
        flags = CONTENT | (busy[id] << 3);
        uint32_t local = (uint32_t)id & 63u;
        uint32_t bits = (1u << (local & 3u)) | (16u << ((local >> 2) & 3u)) | (256u << (local >> 4));
        activity[(size_t)brickCount() + (id >> 6)] |= bits;
      }
      if (expands[id]) flags |= EXPANDS;
      activity[id] |= flags;
    }
  }

  // ---- coarse smoke (smoke.wgsl) --------------------------------------------------------
  // With smokeDivisor 2 or 4 smoke density lives on its own grid, smokeDivisor fine cells
  // wide, in pools of (B / smokeDivisor)^3 cells per slot. Fine evolution stages each
  // step's smoke production in the fields' X channel; the coarse grid transports the
  // density with the same velocity, adds the production's box average and decays it.
  bool coarse() const { return smokeDiv > 1; }
  R smokeCell(const std::vector<R> &buf, int x, int y, int z) const {
    int s = slotAtBrick(x >> sS, y >> sS, z >> sS);
    if (s < 0) return 0;
    int m = bs - 1;
    return buf[(size_t)s * nS + (((z & m) * bs + (y & m)) * bs + (x & m))];
  }
  // A trilinear sample of a smoke pool at `pos` in smoke cells (sampleAt,
  // fluid-sampling.wgsl:85-119), with the range of the eight cells it weighs: the minimum
  // from 1e30 and the maximum from 0, as smokeDonorBounds (smoke.wgsl:59-73) scans them.
  R sampleSmoke(const std::vector<R> &buf, const R pos[3], R *lo = nullptr, R *hi = nullptr) const {
    R q[3], f[3];
    int base[3];
    for (int a = 0; a < 3; a++) {
      q[a] = pos[a] - R(0.5);
      R fl = ffloor(q[a]);
      base[a] = (int)fl;
      f[a] = q[a] - fl;
    }
    R out = 0, mn = R(1e30), mx = 0;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          R wx = x ? f[0] : R(1) - f[0], wy = y ? f[1] : R(1) - f[1], wz = z ? f[2] : R(1) - f[2];
          R v = smokeCell(buf, base[0] + x, base[1] + y, base[2] + z);
          out += v * (wx * wy * wz);
          mn = fmin(mn, v);
          mx = fmax(mx, v);
        }
    if (lo) *lo = mn;
    if (hi) *hi = mx;
    return out;
  }
  void centerS(const int id[3], R w[3]) const {
    for (int a = 0; a < 3; a++) w[a] = ((R)id[a] + R(0.5)) * P.hs;
  }
  template <class Fn> void forActiveSmokeCells(Fn fn) {
    const long count = (long)activeList.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (long i = 0; i < count; i++) {
      const int id = activeList[(size_t)i];
      int bk[3];
      brickOf(id, bk);
      int s = pageSlot(entries[id] & 0xffffff);
      for (int l = 0; l < nS; l++) {
        int c[3] = {bk[0] * bs + (l & (bs - 1)), bk[1] * bs + ((l >> sS) & (bs - 1)), bk[2] * bs + (l >> (2 * sS))};
        fn(id, s, l, c);
      }
    }
  }
  // advectSmoke (smoke.wgsl:104-126): the backward trace from each smoke cell center, the
  // predictor and, for the corrector, the range of its eight donors.
  void advectSmoke() {
    forActiveSmokeCells([&](int, int s, int l, const int c[3]) {
      R pos[3];
      centerS(c, pos);
      R density = 0, lo = 0, hi = 0;
      if (!solidAt(pos)) {
        R cv[3], dep[3], sp[3];
        traceCenter(pos, cv);
        traceBack(pos, cv, dep);
        for (int a = 0; a < 3; a++) sp[a] = dep[a] * P.invHs;
        density = sampleSmoke(S, sp, &lo, &hi);
      }
      Sp[(size_t)s * nS + l] = density;
      SLo[(size_t)s * nS + l] = lo;
      SHi[(size_t)s * nS + l] = hi;
    });
  }
  // The box average of a field pool over the `ratio`^3 fine cells under cell `id` of a grid
  // `ratio` fine cells wide: one trilinear sample at the shared corner of each 2^3 block
  // (velocityCellAverage, fluid-sampling.wgsl:270-286; smokeProduction, smoke.wgsl:24-37).
  void boxAverage4(const std::vector<R> &buf, int ratio, const int id[3], R out[4]) const {
    if (ratio == 1) {
      fieldCell(buf, id[0], id[1], id[2], out);
      return;
    }
    int blocks = ratio / 2;
    R r = (R)ratio;
    R sum[4] = {0, 0, 0, 0};
    for (int z = 0; z < blocks; z++)
      for (int y = 0; y < blocks; y++)
        for (int x = 0; x < blocks; x++) {
          R corner[3] = {(R)id[0] * r + R(1) + R(2) * (R)x, (R)id[1] * r + R(1) + R(2) * (R)y,
                         (R)id[2] * r + R(1) + R(2) * (R)z};
          R v[4];
          sampleField(buf, corner, v);
          for (int k = 0; k < 4; k++) sum[k] += v[k];
        }
    if (blocks == 1) {
      for (int k = 0; k < 4; k++) out[k] = sum[k];
      return;
    }
    R n = (R)(blocks * blocks * blocks);
    for (int k = 0; k < 4; k++) out[k] = sum[k] / n;
  }
  // evolveSmoke (smoke.wgsl:128-171): the MacCormack correction bounded by the donors, then
  // decay and the box average of the fine production (already decayed over half a step),
  // the dormant cleanup, and content with its support on the smoke grid.
  void evolveSmoke() {
    std::vector<uint32_t> busy(brickCount(), 0u);
    const R loss = fexp(-P.dissipation * P.dt);
    forActiveSmokeCells([&](int bid, int s, int l, const int c[3]) {
      R pos[3];
      centerS(c, pos);
      R density = 0;
      if (!solidAt(pos)) {
        R cv[3];
        traceCenter(pos, cv);
        const size_t at = (size_t)s * nS + l;
        R predictor = Sp[at];
        R bounded = predictor;
        if (maccormack) {
          R lo = SLo[at], hi = SHi[at];
          if (lo != hi || predictor != lo) {
            R fw[3], sp[3];
            traceForward(pos, cv, fw);
            for (int a = 0; a < 3; a++) sp[a] = fw[a] * P.invHs;
            R reversed = sampleSmoke(Sp, sp);
            R corrected = predictor + R(0.5) * (S[at] - reversed);
            bounded = (corrected < lo || corrected > hi) ? predictor : corrected;
          }
        }
        R produced[4];
        boxAverage4(F, smokeDiv, c, produced);
        density = fmax(R(0), bounded * loss + produced[0]);
        if (produced[0] == R(0) && produced[2] == R(0) && fmax(density, fmax(produced[1], produced[3])) < R(0.001))
          density = 0;
        if (density > P.cutoff) busy[bid] |= supportOn(bs, P.hv / P.hs, P.invHs, c, cv);
      }
      Sn[(size_t)s * nS + l] = density;
    });
    for (int id : activeList) {
      if (!busy[id]) continue;
      uint32_t local = (uint32_t)id & 63u;
      uint32_t bits = (1u << (local & 3u)) | (16u << ((local >> 2) & 3u)) | (256u << (local >> 4));
      activity[(size_t)brickCount() + (id >> 6)] |= bits;
      activity[id] |= CONTENT | (busy[id] << 3);
    }
  }
  // smokeVelocityAverage (smoke.wgsl:4-21): the smoke density under velocity cell `p`, for
  // buoyancy. With velocity cells at most as wide as smoke cells, one sample at the cell's
  // center; with velocity cells two smoke cells wide, their box average.
  R smokeVelocityAverage(const int p[3]) const {
    R sp[3];
    if (smokeRatio == 1) {
      R w[3];
      centerV(p, w);
      for (int a = 0; a < 3; a++) sp[a] = w[a] * P.invHs;
    } else {
      for (int a = 0; a < 3; a++) sp[a] = (R)p[a] * R(2) + R(1);
    }
    return sampleSmoke(S, sp);
  }

  // ---- forces and divergence (fluid-forces.wgsl) --------------------------------------
  R curlCentered(int x, int y, int z, int comp) const {
    int bx = x, by = y, bz = z;
    if (comp == 0) bx--;
    if (comp == 1) by--;
    if (comp == 2) bz--;
    R a = velCell(Vcor, comp, x, y, z), bb = velCell(Vcor, comp, bx, by, bz);
    return R(0.5) * (a + bb);
  }
  void computeCurl() {
    forActiveVelocityCells([&](int s, int l, const int p[3]) {
      R hx = R(0.5) / P.hv;
      int x = p[0], y = p[1], z = p[2];
      R dx0 = (curlCentered(x + 1, y, z, 1) - curlCentered(x - 1, y, z, 1)) * hx;
      R dx1 = (curlCentered(x + 1, y, z, 2) - curlCentered(x - 1, y, z, 2)) * hx;
      R dy0 = (curlCentered(x, y + 1, z, 0) - curlCentered(x, y - 1, z, 0)) * hx;
      R dy1 = (curlCentered(x, y + 1, z, 2) - curlCentered(x, y - 1, z, 2)) * hx;
      R dz0 = (curlCentered(x, y, z + 1, 0) - curlCentered(x, y, z - 1, 0)) * hx;
      R dz1 = (curlCentered(x, y, z + 1, 1) - curlCentered(x, y, z - 1, 1)) * hx;
      R c0 = dy1 - dz1, c1 = dz0 - dx1, c2 = dx0 - dy0;
      R *o = &Curl[((size_t)s * nV + l) * 4];
      o[0] = c0;
      o[1] = c1;
      o[2] = c2;
      o[3] = length3(c0, c1, c2);
    });
  }
  // curlNoise (fluid-forces.wgsl:3-9): the table sampled trilinearly with wrapping.
  void curlNoise(const R p[3], R out[3]) const {
    // rotation columns (.36,-.8,.48), (.48,.6,.64), (-.8,0,.6)
    const R c0[3] = {R(0.36), R(-0.8), R(0.48)}, c1[3] = {R(0.48), R(0.6), R(0.64)}, c2[3] = {R(-0.8), R(0), R(0.6)};
    R u[3];
    for (int a = 0; a < 3; a++) u[a] = (c0[a] * p[0] + c1[a] * p[1] + c2[a] * p[2]) / R(16);
    R s[4] = {0, 0, 0, 0};
    R f[3];
    int i0[3];
    for (int a = 0; a < 3; a++) {
      R t = u[a] * R(64) - R(0.5);
      R fl = ffloor(t);
      i0[a] = (int)fl;
      f[a] = t - fl;
    }
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          R w = (x ? f[0] : R(1) - f[0]) * (y ? f[1] : R(1) - f[1]) * (z ? f[2] : R(1) - f[2]);
          int ix = (i0[0] + x) & 63, iy = (i0[1] + y) & 63, iz = (i0[2] + z) & 63;
          const float *t = &noise[(((size_t)iz * 64 + iy) * 64 + ix) * 4];
          for (int k = 0; k < 3; k++) s[k] += (R)t[k] * w;
        }
    for (int a = 0; a < 3; a++) out[a] = (a == 0 ? c0 : a == 1 ? c1 : c2)[0] * s[0] +
                                         (a == 0 ? c0 : a == 1 ? c1 : c2)[1] * s[1] +
                                         (a == 0 ? c0 : a == 1 ? c1 : c2)[2] * s[2];
  }
  // coverage + externalForces (force-fields.wgsl:50-103).
  void externalForces(const int cell[3], const R p[3], R out[3]) const {
    out[0] = out[1] = out[2] = 0;
    for (const Force<R> &f : forces) {
      R weight = 1;
      if (f.targeted && f.meshTargets && f.covered.count({cell[0], cell[1], cell[2]})) {
        weight = 1;  // meshCovers (force-fields.wgsl:27-58)
      } else if (f.targeted) {
        R covered = 0;
        for (const Region<R> &r : f.regions) {
          R dd[3];
          for (int a = 0; a < 3; a++) dd[a] = fabs_((p[a] - r.center[a]) * r.invSize[a]);
          R dist = r.kind == R(3) ? fmax(dd[0], fmax(dd[1], dd[2])) : length3(dd[0], dd[1], dd[2]);
          covered = fmax(covered, R(1) - smoothstepv(R(0.5), R(1), dist));
        }
        weight = covered;
      }
      if (weight == R(0)) continue;
      R acc[3] = {f.axis[0] * f.strength, f.axis[1] * f.strength, f.axis[2] * f.strength};
      if (f.kind == R(1)) {
        R q[3] = {p[0] * f.scale * R(0.65) + P.seed, p[1] * f.scale * R(0.65) + -P.time * R(0.7),
                  p[2] * f.scale * R(0.65) + P.seed * R(0.25)};
        R n1[3], n2[3];
        curlNoise(q, n1);
        R q2[3] = {q[0] * R(2.03) + R(13.7), q[1] * R(2.03) + R(5.3), q[2] * R(2.03) + R(19.1)};
        curlNoise(q2, n2);
        for (int a = 0; a < 3; a++) acc[a] = (n1[a] + R(0.45) * n2[a]) * f.strength;
      } else if (f.kind >= R(2)) {
        R delta[3] = {p[0] - f.center[0], p[1] - f.center[1], p[2] - f.center[2]};
        R radial[3] = {delta[0], delta[1], delta[2]};
        if (f.kind == R(2)) {
          R dp = delta[0] * f.axis[0] + delta[1] * f.axis[1] + delta[2] * f.axis[2];
          for (int a = 0; a < 3; a++) radial[a] = radial[a] - f.axis[a] * dp;
        }
        R distance = length3(radial[0], radial[1], radial[2]);
        R radius = fmax(f.scale, R(0.01));
        R dir_[3];
        for (int a = 0; a < 3; a++) dir_[a] = radial[a] / fmax(distance, radius * R(0.25));
        R rr = radial[0] * radial[0] + radial[1] * radial[1] + radial[2] * radial[2];
        R falloff = R(1) / (R(1) + rr / (radius * radius));
        if (f.kind == R(2)) {
          R cr[3] = {f.axis[1] * dir_[2] - f.axis[2] * dir_[1], f.axis[2] * dir_[0] - f.axis[0] * dir_[2],
                     f.axis[0] * dir_[1] - f.axis[1] * dir_[0]};
          for (int a = 0; a < 3; a++)
            acc[a] = (cr[a] * f.strength - dir_[a] * f.inward + f.axis[a] * f.lift) * falloff;
        } else {
          for (int a = 0; a < 3; a++) acc[a] = dir_[a] * f.strength * falloff;
        }
      }
      for (int a = 0; a < 3; a++) out[a] += acc[a] * weight;
    }
  }
  // velocityCellAverage (fluid-sampling.wgsl:270-286) of the field grid under cell `id`.
  void velocityCellAverage4(const int id[3], R out[4]) const { boxAverage4(F, d, id, out); }
  R expansionAverage(const int id[3]) const {
    if (d == 1) return fieldScalar(E, id[0], id[1], id[2]);
    int blocks = d / 2;
    R ratio = (R)d;
    R sum = 0;
    for (int z = 0; z < blocks; z++)
      for (int y = 0; y < blocks; y++)
        for (int x = 0; x < blocks; x++) {
          R corner[3] = {(R)id[0] * ratio + R(1) + R(2) * (R)x, (R)id[1] * ratio + R(1) + R(2) * (R)y,
                         (R)id[2] * ratio + R(1) + R(2) * (R)z};
          // sampleCells on the expansion pool
          R q[3], f[3];
          int base[3];
          for (int a = 0; a < 3; a++) {
            q[a] = corner[a] - R(0.5);
            R fl = ffloor(q[a]);
            base[a] = (int)fl;
            f[a] = q[a] - fl;
          }
          R v = 0;
          for (int zz = 0; zz < 2; zz++)
            for (int yy = 0; yy < 2; yy++)
              for (int xx = 0; xx < 2; xx++) {
                R w = (xx ? f[0] : R(1) - f[0]) * (yy ? f[1] : R(1) - f[1]) * (zz ? f[2] : R(1) - f[2]);
                v += fieldScalar(E, base[0] + xx, base[1] + yy, base[2] + zz) * w;
              }
          sum += v;
        }
    return blocks == 1 ? sum : sum / (R)(blocks * blocks * blocks);
  }
  // centerForce (fluid-forces.wgsl:79-104).
  void centerForce(const int p[3], R out[3]) const {
    R w[3];
    centerV(p, w);
    if (solidAt(w)) {
      out[0] = out[1] = out[2] = 0;
      return;
    }
    R f[4];
    velocityCellAverage4(p, f);
    R conf[3] = {0, 0, 0};
    if (P.vorticity != R(0)) {
      R c[3] = {curlCell(0, p[0], p[1], p[2]), curlCell(1, p[0], p[1], p[2]), curlCell(2, p[0], p[1], p[2])};
      R eta[3] = {curlCell(3, p[0] + 1, p[1], p[2]) - curlCell(3, p[0] - 1, p[1], p[2]),
                  curlCell(3, p[0], p[1] + 1, p[2]) - curlCell(3, p[0], p[1] - 1, p[2]),
                  curlCell(3, p[0], p[1], p[2] + 1) - curlCell(3, p[0], p[1], p[2] - 1)};
      R len = length3(eta[0], eta[1], eta[2]) + R(0.0001);
      R n[3] = {eta[0] / len, eta[1] / len, eta[2] / len};
      R cr[3] = {n[1] * c[2] - n[2] * c[1], n[2] * c[0] - n[0] * c[2], n[0] * c[1] - n[1] * c[0]};
      for (int a = 0; a < 3; a++) conf[a] = cr[a] * P.vorticity * P.hv;
    }
    R ext[3];
    externalForces(p, w, ext);
    out[0] = conf[0] + ext[0] + R(0);
    const R density = coarse() ? smokeVelocityAverage(p) : f[0];
    out[1] = conf[1] + ext[1] + (P.buoyancy * f[1] - P.smokeWeight * density);
    out[2] = conf[2] + ext[2] + R(0);
  }
  void cellForce(const int p[3], R out[3]) const {
    if (!computes(p[0] >> sV, p[1] >> sV, p[2] >> sV)) {
      out[0] = out[1] = out[2] = 0;
      return;
    }
    centerForce(p, out);
  }
  // sourceVelocity (fluid-forces.wgsl:122-149).
  R sourceVelocity(R value, const R face[3], int comp) const {
    R result = value;
    for (const Emitter<R> &e : emitters) {
      if (e.response == R(0)) continue;
      R wgt = emitterWeight(e, face);
      if (wgt == R(0)) continue;
      R radial = 0;
      if (e.outward != R(0)) {
        R dd[3] = {face[0] - e.pos[0], face[1] - e.pos[1], face[2] - e.pos[2]};
        radial = dd[comp] / fmax(length3(dd[0], dd[1], dd[2]), R(1e-6)) * e.outward;
      }
      result = mixv(result, e.vel[comp] + radial, R(1) - fexp(-e.response * wgt * P.dt));
    }
    if (meshCount > 0) {
      const R *motion = &meshRecords[8 * (size_t)meshRecordAt(face) + 4];
      if (motion[3] > R(0)) result = mixv(result, motion[comp] / motion[3], R(1) - fexp(-motion[3] * P.dt));
    }
    return result;
  }
  void applyForces() {
    R damp = fexp(-P.damping * P.dt);
    forActiveVelocityCells([&](int s, int l, const int p[3]) {
      R fc[3], fx[3], fy[3], fz[3];
      cellForce(p, fc);
      int px[3] = {p[0] + 1, p[1], p[2]}, py[3] = {p[0], p[1] + 1, p[2]}, pz[3] = {p[0], p[1], p[2] + 1};
      cellForce(px, fx);
      cellForce(py, fy);
      cellForce(pz, fz);
      R force[3] = {R(0.5) * (fc[0] + fx[0]), R(0.5) * (fc[1] + fy[1]), R(0.5) * (fc[2] + fz[2])};
      R pos[3];
      centerV(p, pos);
      const R *vc = &Vcor[(size_t)s * nV * 3];
      R v[3];
      for (int a = 0; a < 3; a++) v[a] = vc[a * nV + l] * damp + P.dt * force[a];
      R half = R(0.5) * P.hv;
      R out[3];
      for (int a = 0; a < 3; a++) {
        R face[3] = {pos[0], pos[1], pos[2]};
        face[a] = pos[a] + half;
        out[a] = sourceVelocity(v[a], face, a);
      }
      // enforceFaces (fluid-forces.wgsl:15-31)
      R *o = &Vf[(size_t)s * nV * 3];
      if (solidAt(pos)) {
        o[l] = o[nV + l] = o[2 * nV + l] = 0;
        return;
      }
      for (int a = 0; a < 3; a++) {
        R q[3] = {pos[0], pos[1], pos[2]};
        q[a] = pos[a] + P.hv;
        if (solidAt(q)) out[a] = 0;
        o[a * nV + l] = out[a];
      }
    });
  }
  void computeDivergence() {
    forActiveVelocityCells([&](int s, int l, const int p[3]) {
      R h = P.hv;
      const R *vf = &Vf[(size_t)s * nV * 3];
      R div = (vf[l] - velCell(Vf, 0, p[0] - 1, p[1], p[2])) * P.invHv +
              (vf[nV + l] - velCell(Vf, 1, p[0], p[1] - 1, p[2])) * P.invHv +
              (vf[2 * nV + l] - velCell(Vf, 2, p[0], p[1], p[2] - 1)) * P.invHv;
      D[(size_t)s * nV + l] = h * h * (div - expansionAverage(p));
    });
  }

  // ---- multigrid (PressureSolver.ts:276-402, pressure.wgsl) ---------------------------
  int levelShift(int L) const { return L < brickLevels ? sV - L : sV + 2 - L; }
  // cellAddress (pressure.wgsl:71-92): owner slot/tile and local index, or -1.
  long addr(const int p[3], int L) const {
    int sh = levelShift(L), sd = 1 << sh;
    int ox = p[0] >> sh, oy = p[1] >> sh, oz = p[2] >> sh;
    int owner;
    if (L < brickLevels) {
      owner = slotAtBrick(ox, oy, oz);
    } else {
      owner = dir.slotAt(ox, oy, oz);
    }
    if (owner < 0) return -1;
    int lx = p[0] & (sd - 1), ly = p[1] & (sd - 1), lz = p[2] & (sd - 1);
    return (long)owner * sd * sd * sd + (lz * sd + ly) * sd + lx;
  }
  // activeCell (pressure.wgsl:101-132).
  bool activeCell(const int p[3], int L) const {
    if (L >= brickLevels) {
      int n = 1 << (L - brickLevels);
      for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
          for (int x = 0; x < n; x++) {
            int id = brickIdOf(p[0] * n + x, p[1] * n + y, p[2] * n + z);
            if (id < 0 || entries[id] < 0 || records[(size_t)id * BRICK_RECORD_WORDS + PRESSURE_FLAG] == 0)
              return false;
          }
      return true;
    }
    int span = 1 << L;
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      lo[a] = (p[a] * span - 1) >> sV;
      hi[a] = (p[a] * span + span - 1) >> sV;
    }
    if (P.ground) lo[1] = std::max(lo[1], 0);
    for (int z = lo[2]; z <= hi[2]; z++)
      for (int y = lo[1]; y <= hi[1]; y++)
        for (int x = lo[0]; x <= hi[0]; x++)
          if (!computes(x, y, z)) return false;
    return true;
  }
  bool blockedAt(const int p[3], int L) const {
    R scale = (R)(1 << L);
    R w[3];
    for (int a = 0; a < 3; a++) w[a] = ((R)p[a] + R(0.5)) * P.hv * scale;
    return solidAt(w);
  }
  uint8_t cellTopology(const int p[3], int L) const {
    if (blockedAt(p, L)) return STORED_SOLID;
    static const int OFF[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    uint8_t flags = 0;
    for (int sdx = 0; sdx < 6; sdx++) {
      int q[3] = {p[0] + OFF[sdx][0], p[1] + OFF[sdx][1], p[2] + OFF[sdx][2]};
      if (!blockedAt(q, L)) flags |= (uint8_t)(1u << sdx);
    }
    return flags;
  }
  // buildBrickTopology (pressure.wgsl:259-278) for one brick, all brick levels.
  void brickTopology(int id) {
    int e = entries[id];
    if (e < 0) return;
    int slot = pageSlot(e & 0xffffff);
    int bk[3];
    brickOf(id, bk);
    for (int L = 0; L < brickLevels; L++) {
      int sd = levels[L].side;
      for (int l = 0; l < sd * sd * sd; l++) {
        int p[3] = {bk[0] * sd + l % sd, bk[1] * sd + (l / sd) % sd, bk[2] * sd + l / (sd * sd)};
        levels[L].topo[(size_t)slot * sd * sd * sd + l] = cellTopology(p, L);
      }
    }
  }
  void tileTopology() {
    for (int L = brickLevels; L <= bottom; L++) {
      int sd = levels[L].side;
      for (int t = 0; t < capacity; t++) {
        const int32_t *r = &tables.records[(size_t)t * TILE_RECORD_WORDS];
        for (int l = 0; l < sd * sd * sd; l++) {
          uint8_t v = 0;
          if (r[3] != TILE_FREE) {
            int p[3] = {r[0] * sd + l % sd, r[1] * sd + (l / sd) % sd, r[2] * sd + l / (sd * sd)};
            v = cellTopology(p, L);
          }
          levels[L].topo[(size_t)t * sd * sd * sd + l] = v;
        }
      }
    }
  }

  // The cells a level's sweep visits: every cell of every slotted brick on the brick
  // levels, of every live or closing tile on the tile levels. With `parallel` the bricks
  // (tiles) run on the host's threads; a pass that sums over cells passes false.
  template <class Fn> void forLevelCells(int L, Fn fn, bool parallel = true) {
    int sd = levels[L].side;
    if (L < brickLevels) {
      const long count = (long)activeList.size();
#pragma omp parallel for schedule(dynamic, 8) if (parallel)
      for (long i = 0; i < count; i++) {
        const int id = activeList[(size_t)i];
        int bk[3];
        brickOf(id, bk);
        for (int l = 0; l < sd * sd * sd; l++) {
          int p[3] = {bk[0] * sd + l % sd, bk[1] * sd + (l / sd) % sd, bk[2] * sd + l / (sd * sd)};
          fn(p);
        }
      }
      return;
    }
#pragma omp parallel for schedule(dynamic, 8) if (parallel)
    for (int t = 0; t < capacity; t++) {
      const int32_t *r = &tables.records[(size_t)t * TILE_RECORD_WORDS];
      if (r[3] == TILE_FREE) continue;
      for (int l = 0; l < sd * sd * sd; l++) {
        int p[3] = {r[0] * sd + l % sd, r[1] * sd + (l / sd) % sd, r[2] * sd + l / (sd * sd)};
        fn(p);
      }
    }
  }
  R readAt(const std::vector<R> &v, const int p[3], int L) const {
    long a = addr(p, L);
    return a < 0 ? R(0) : v[a];
  }
  int flagsOf(const int p[3], int L) const {  // expandFlags: 1024 = solid
    long a = addr(p, L);
    if (a < 0) return 1024;
    uint8_t s = levels[L].topo[a];
    return (s & STORED_SOLID) ? 1024 : (s & 63);
  }
  R axisWeight() const {
    R inv = R(1) / P.hv;
    R ref = R(1) / inv;
    R nrm = ref * inv;
    return nrm * nrm;
  }
  static R openCount(int flags, int sdx) { return (R)((flags >> sdx) & 1) + (R)((flags >> (sdx + 1)) & 1); }
  // stencil (pressure.wgsl:376-402): neighbor sum over open sides, and the diagonal.
  void stencil(const std::vector<R> &pv, const int p[3], int flags, int L, R &sum, R &diag) const {
    static const int OFF[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    R n[6];
    for (int sdx = 0; sdx < 6; sdx++) {
      if (!((flags >> sdx) & 1)) {
        n[sdx] = 0;
        continue;
      }
      int q[3] = {p[0] + OFF[sdx][0], p[1] + OFF[sdx][1], p[2] + OFF[sdx][2]};
      n[sdx] = readAt(pv, q, L);
    }
    R w = axisWeight();
    sum = w * (n[0] + n[1]) + w * (n[2] + n[3]) + w * (n[4] + n[5]);
    diag = w * openCount(flags, 0) + w * openCount(flags, 2) + w * openCount(flags, 4);
  }
  // relaxCell (pressure.wgsl:468-479), ping-pong: reads lv.p, writes lv.q.
  void relax(int L) {
    Level &lv = levels[L];
    forLevelCells(L, [&](const int p[3]) {
      long a = addr(p, L);
      int flags = flagsOf(p, L);
      if ((flags & 1024) || !activeCell(p, L)) {
        lv.q[a] = 0;
        return;
      }
      R sum, diag;
      stencil(lv.p, p, flags, L, sum, diag);
      R cand = (sum - lv.rhs[a]) / fmax(diag, R(1e-12));
      lv.q[a] = stored(mixv(lv.p[a], cand, R(0.8)));
    });
    std::swap(lv.p, lv.q);
  }
  // initializePressure (pressure.wgsl:485-505): the first iterate from zero.
  void initialize0() {
    Level &lv = levels[0];
    forLevelCells(0, [&](const int p[3]) {
      long a = addr(p, 0);
      int flags = flagsOf(p, 0);
      R first = 0;
      if (!(flags & 1024) && activeCell(p, 0)) {
        R w = axisWeight();
        R diag = w * openCount(flags, 0) + w * openCount(flags, 2) + w * openCount(flags, 4);
        first = mixv(R(0), (R(0) - lv.rhs[a]) / fmax(diag, R(1e-12)), R(0.8));
      }
      lv.p[a] = stored(first);
    });
  }
  // restrictCell (pressure.wgsl:601-647): coarse rhs and its first iterate.
  void restrictTo(int L) {
    Level &fine = levels[L], &coarse = levels[L + 1];
    forLevelCells(L + 1, [&](const int id[3]) {
      R residual = 0;
      bool every = true;
      for (int z = 0; z < 2; z++)
        for (int y = 0; y < 2; y++)
          for (int x = 0; x < 2; x++) {
            int p[3] = {id[0] * 2 + x, id[1] * 2 + y, id[2] * 2 + z};
            bool unknown = activeCell(p, L);
            every = every && unknown;
            if (!unknown) continue;
            int flags = flagsOf(p, L);
            if (flags & 1024) continue;
            long a = addr(p, L);
            R sum, diag;
            stencil(fine.p, p, flags, L, sum, diag);
            residual += fine.rhs[a] - (sum - diag * fine.p[a]);
          }
      R rhs = residual * R(0.5);
      long out = addr(id, L + 1);
      coarse.rhs[out] = stored(rhs);
      R first = 0;
      if (every) {
        int flags = flagsOf(id, L + 1);
        if (!(flags & 1024)) {
          R w = axisWeight();
          R diag = w * openCount(flags, 0) + w * openCount(flags, 2) + w * openCount(flags, 4);
          first = mixv(R(0), (R(0) - rhs) / fmax(diag, R(1e-12)), R(0.8));
        }
      }
      coarse.p[out] = stored(first);
    });
  }
  bool coarseSolid(const int c[3], int L) const {
    long a = addr(c, L);
    if (a < 0) return false;
    return (levels[L].topo[a] & STORED_SOLID) != 0;
  }
  // coarseValue (pressure.wgsl:670-691): Neumann ghosts beside solids, a mirror below the floor.
  R coarseValue(const int p[3], int L) const {
    const Level &cv = levels[L];
    int groundY = P.ground ? 0 : INT32_MIN;
    int c[3] = {p[0], std::max(p[1], groundY), p[2]};
    if (!coarseSolid(c, L)) return readAt(cv.p, c, L);
    R sum = 0, count = 0;
    for (int axis = 0; axis < 3; axis++)
      for (int sd = -1; sd <= 1; sd += 2) {
        int q[3] = {c[0], c[1], c[2]};
        q[axis] += sd;
        q[1] = std::max(q[1], groundY);
        if (!coarseSolid(q, L)) {
          sum += readAt(cv.p, q, L);
          count += R(1);
        }
      }
    return sum / fmax(count, R(1));
  }
  // prolongateCell (pressure.wgsl:726-755).
  void prolong(int L) {
    Level &lv = levels[L];
    const Level &cv = levels[L + 1];
    forLevelCells(L, [&](const int p[3]) {
      long a = addr(p, L);
      int flags = flagsOf(p, L);
      if ((flags & 1024) || !activeCell(p, L)) {
        lv.q[a] = 0;
        return;
      }
      R q[3], f[3];
      int i[3];
      for (int k = 0; k < 3; k++) {
        q[k] = ((R)p[k] + R(0.5)) * R(0.5) - R(0.5);
        R fl = ffloor(q[k]);
        i[k] = (int)fl;
        f[k] = q[k] - fl;
      }
      bool plain = P.colliders == 0 && (!P.ground || i[1] >= 0);
      auto at = [&](int dx, int dy, int dz) {
        int c[3] = {i[0] + dx, i[1] + dy, i[2] + dz};
        return plain ? readAt(cv.p, c, L + 1) : coarseValue(c, L + 1);
      };
      R corr = mixv(mixv(mixv(at(0, 0, 0), at(1, 0, 0), f[0]), mixv(at(0, 1, 0), at(1, 1, 0), f[0]), f[1]),
                    mixv(mixv(at(0, 0, 1), at(1, 0, 1), f[0]), mixv(at(0, 1, 1), at(1, 1, 1), f[0]), f[1]), f[2]);
      lv.q[a] = stored(lv.p[a] + corr);
    });
    std::swap(lv.p, lv.q);
  }
  // The coarsest level's weighted Jacobi sweeps in FP64 (one cell per tile slot), the
  // result rounded to FP32 for the float restatement, then stored.
  void relaxBottom64(int iterations) {
    Level &lv = levels[bottom];
    size_t n = lv.p.size();
    std::vector<double> val(n, 0.0), next(n, 0.0);
    std::vector<char> live(n, 0);
    std::vector<int> flg(n, 0);
    forLevelCells(bottom, [&](const int p[3]) {
      long a = addr(p, bottom);
      int flags = flagsOf(p, bottom);
      bool unknown = activeCell(p, bottom);
      live[a] = unknown && !(flags & 1024);
      flg[a] = flags & 63;
      val[a] = unknown ? (double)lv.p[a] : 0.0;
    });
    const double w = (double)axisWeight();
    static const int OFF[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int it = 0; it < iterations; it++) {
      forLevelCells(bottom, [&](const int p[3]) {
        long a = addr(p, bottom);
        if (!live[a]) {
          next[a] = 0.0;
          return;
        }
        double nb[6];
        for (int sd = 0; sd < 6; sd++) {
          nb[sd] = 0.0;
          if (!((flg[a] >> sd) & 1)) continue;
          int q[3] = {p[0] + OFF[sd][0], p[1] + OFF[sd][1], p[2] + OFF[sd][2]};
          long b = addr(q, bottom);
          if (b >= 0) nb[sd] = val[b];
        }
        double sum = w * (nb[0] + nb[1]) + w * (nb[2] + nb[3]) + w * (nb[4] + nb[5]);
        double diag = w * (double)openCount(flg[a], 0) + w * (double)openCount(flg[a], 2) +
                      w * (double)openCount(flg[a], 4);
        double cand = (sum - (double)lv.rhs[a]) / (diag < 1e-12 ? 1e-12 : diag);
        next[a] = val[a] * (1.0 - 0.8) + cand * 0.8;
      });
      forLevelCells(bottom, [&](const int p[3]) {
        long a = addr(p, bottom);
        val[a] = next[a];
      });
    }
    forLevelCells(bottom, [&](const int p[3]) {
      long a = addr(p, bottom);
      if constexpr (std::is_same<R, float>::value) lv.p[a] = stored((float)val[a]);
      else lv.p[a] = (R)val[a];
    });
  }
  void cycle(int L) {
    if (L == bottom) {
      if (pressureMode != 0) relaxBottom64(BOTTOM_ITERATIONS);
      else
        for (int i = 0; i < BOTTOM_ITERATIONS; i++) relax(L);
      return;
    }
    if (L == 0) initialize0();
    else {
      relax(L);
      relax(L);
    }
    restrictTo(L);
    cycle(L + 1);
    prolong(L);
    relax(L);
    relax(L);
    relax(L);
  }
  void solvePressure() {
    // Level 0's right-hand side is the divergence, in slot order.
    levels[0].rhs = D;
    if (geometryDirty) {
      for (int id = 0; id < brickCount(); id++)
        if (tileState(id >> 6) != TILE_FREE) brickTopology(id);
      tileTopology();
      geometryDirty = false;
    }
    // Fresh slots' topology, every step (PressureSolver.ts:291-302).
    for (auto &z : zeroList) brickTopology(z.first);
    if (pressureMode == 0) {
      cycle(0);
      Pfine = levels[0].p;
      PfineLo.clear();
      lastCycles = 1;
      lastResidual = -1;
      return;
    }
    correct();
    Pfine = dcX;
    PfineLo = dcXlo;
  }
  // hi + lo += v: Knuth's two-sum of hi and v, its error added to lo, then renormalized so
  // that hi is the rounded value of hi + lo (no counterpart in the reference).
  static void addToPair(R &hi, R &lo, R v) {
    const R s = hi + v;
    const R bp = s - hi;
    const R err = (hi - (s - bp)) + (v - bp);
    const R l = lo + err;
    const R h = s + l;
    lo = l - (h - s);
    hi = h;
  }
  // Defect correction: x = 0; r = b - A x (sums of r^2 and b^2 in FP64); while
  // |r| / |b| > tol, one V-cycle on A e = r / s, s the largest power of two not above
  // max |r| (1 for zero or subnormal residuals), and x += s e.
  void correct() {
    const std::vector<R> b = D;
    dcX.assign(b.size(), R(0));
    dcXlo.assign(b.size(), R(0));
    std::vector<R> r(b.size(), R(0));
    lastCycles = 0;
    lastResidual = 1;
    for (int it = 0; it <= pressureMaxCycles; it++) {
      double rr = 0, bb = 0;
      R big = 0;
      forLevelCells(0, [&](const int p[3]) {
        long a = addr(p, 0);
        int flags = flagsOf(p, 0);
        R v = 0;
        if (!(flags & 1024) && activeCell(p, 0)) {
          // The row's dot product in FP64 from x = hi + lo, rounded once.
          static const int OFF[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
          const double w = (double)axisWeight();
          double sum = 0;
          for (int k = 0; k < 6; k++) {
            if (!((flags >> k) & 1)) continue;
            int q[3] = {p[0] + OFF[k][0], p[1] + OFF[k][1], p[2] + OFF[k][2]};
            sum += w * ((double)readAt(dcX, q, 0) + (double)readAt(dcXlo, q, 0));
          }
          R w1 = axisWeight();
          R diag = w1 * openCount(flags, 0) + w1 * openCount(flags, 2) + w1 * openCount(flags, 4);
          v = (R)((double)b[a] - (sum - (double)diag * ((double)dcX[a] + (double)dcXlo[a])));
          rr += (double)v * (double)v;
          bb += (double)b[a] * (double)b[a];
          big = fmax(big, fabs_(v));
        }
        r[a] = v;
      }, false);
      lastResidual = bb > 0 ? std::sqrt(rr / bb) : 0.0;
      if (lastResidual <= pressureTol || it == pressureMaxCycles) break;
      R scale = 1;
      if constexpr (std::is_same<R, float>::value) {
        uint32_t bits;
        float m = big;
        std::memcpy(&bits, &m, 4);
        bits &= 0x7f800000u;
        if (bits) std::memcpy(&scale, &bits, 4);
      } else {
        int e;
        std::frexp(big, &e);
        scale = big > 0 ? std::ldexp(1.0, e - 1) : 1.0;
      }
      R inv = R(1) / scale;
      Level &l0 = levels[0];
      forLevelCells(0, [&](const int p[3]) {
        long a = addr(p, 0);
        l0.rhs[a] = stored(r[a] * inv);
      });
      cycle(0);
      forLevelCells(0, [&](const int p[3]) {
        long a = addr(p, 0);
        int flags = flagsOf(p, 0);
        R hi = 0, lo = 0;
        if (!(flags & 1024) && activeCell(p, 0)) {
          hi = dcX[a];
          lo = dcXlo[a];
          addToPair(hi, lo, scale * levels[0].p[a]);
        }
        dcX[a] = hi;
        dcXlo[a] = lo;
      });
      lastCycles = it + 1;
    }
  }

  // project (fluid-forces.wgsl:205-262).
  bool pressureUnknown(const int p[3]) const {
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      lo[a] = (p[a] - 1) >> sV;
      hi[a] = p[a] >> sV;
    }
    if (P.ground) lo[1] = std::max(lo[1], 0);
    for (int z = lo[2]; z <= hi[2]; z++)
      for (int y = lo[1]; y <= hi[1]; y++)
        for (int x = lo[0]; x <= hi[0]; x++)
          if (!computes(x, y, z)) return false;
    return true;
  }
  void project() {
    std::vector<R> out(V.size(), R(0));
    out = V;  // slots that no brick holds keep their values
    forActiveVelocityCells([&](int s, int l, const int p[3]) {
      R pos[3];
      centerV(p, pos);
      R *o = &out[(size_t)s * nV * 3];
      if (solidAt(pos)) {
        o[l] = o[nV + l] = o[2 * nV + l] = 0;
        return;
      }
      R cp = velScalar(Pfine, p[0], p[1], p[2]);
      const bool pair = !PfineLo.empty();
      const R cpLo = pair ? velScalar(PfineLo, p[0], p[1], p[2]) : R(0);
      const R *vf = &Vf[(size_t)s * nV * 3];
      bool here = pressureUnknown(p);
      for (int a = 0; a < 3; a++) {
        int nb[3] = {p[0], p[1], p[2]};
        nb[a] += 1;
        R npos[3];
        centerV(nb, npos);
        R value = vf[a * nV + l];
        R res;
        if (solidAt(npos)) res = 0;
        else if (!here && !pressureUnknown(nb)) res = 0;
        else if (!pair) res = value - (velScalar(Pfine, nb[0], nb[1], nb[2]) - cp) * P.invHv;
        else {
          const R g = (velScalar(Pfine, nb[0], nb[1], nb[2]) - cp) + (velScalar(PfineLo, nb[0], nb[1], nb[2]) - cpLo);
          res = value - g * P.invHv;
        }
        o[a * nV + l] = res;
      }
    });
    V.swap(out);
  }

  // Readback and readActivity (FluidSimulation.ts:1514-1523, :1701-1746), applied at once.
  void readActivity() {
    contentTiles.clear();
    for (int t = 0; t < capacity; t++) {
      uint32_t flags = activity[(size_t)brickCount() + t];
      if (flags && t < (int)dir.snapshot.size() && dir.snapshot[t][3])
        contentTiles.push_back({{dir.snapshot[t][0], dir.snapshot[t][1], dir.snapshot[t][2]}, flags});
    }
    lastFailures = failures;
    uint32_t fails = failures;
    failures = 0;
    contentStep = steps;
    activeBrickCount = (int64_t)activeList.size();
    if (steps < poolGrownAt) return;
    usedSlots = poolSlots - freeCount;
    int limit = slotLimit();
    poolLimited = fails > 0 && poolSlots >= limit;
    if (fails > 0 && poolSlots < limit)
      poolTarget = std::max(std::max(poolTarget, (int)std::ceil((usedSlots + (double)fails) * 1.5)),
                            poolSlots + side);
  }

  void reset() {
    allocate();
    time = 0;
    steps = 0;
    mayHaveFlame = false;
    poolGrownAt = 0;
    activeBrickCount = 0;
  }
};

}  // namespace oracle
