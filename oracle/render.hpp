// ORACLE — test infrastructure only. Nothing in the shipped library may include, link or
// call this file; tests, smoke() and bench.py's CPU baseline use it as the checker.
//
// A sequential CPU restatement of Fire Pro's render preparation and camera ray-march over
// the solver's bricks, at single precision (the twin of the HIP render path):
//  * the blackbody table (src/engine/blackbody.ts:5-97);
//  * VolumeLighting.encode's schedule (src/engine/VolumeLighting.ts:249-375) and its
//    kernels (src/engine/shaders/lighting.wgsl): exportPages, smoothSmoke,
//    buildOccupancy, buildLighting/lightSlot/incidentLight, lightFresh;
//  * the camera march (volume-march.wgsl, volume-reconstruction.wgsl,
//    volume-bricks.wgsl, fire-emission.wgsl, scene-volume.wgsl:38-48 and :146-167).
// Logical semantics as in the solver oracle: page tables say which bricks of a box have
// slots; every filtered fetch is an exact trilinear sum in FP32 over the texels the
// reference's hardware filter reads (the slot with its apron when the base brick has a
// slot, cell by cell otherwise). The render field and the incident light keep the
// reference's one-texel aprons, because their apron values are computed, not copied.
// Deliberate differences: no half-precision stores (the reference keeps rgba16float
// render fields, light and blackbody texels); pow(x, 1.5) is x * sqrt(x); exp is fexp
// and exp2 is fexp2; no half-resolution slices, no scene depth and no debug views. With
// coarse smoke, the scene lights' sample at a lattice point reads the logical neighbor
// cells, where the reference's in-slot texture fetch can reach the atlas texel before the
// slot (brick 8 with smokeDivisor 4).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "fmath.hpp"
#include "solver.hpp"

namespace oracle {

constexpr int SHADOW_STEPS = 32;  // src/engine/types.ts SHADOW_STEPS
constexpr int LUT_SIZE = 512;     // blackbody.ts:73

// ---- blackbody table (blackbody.ts:5-97), in double, stored as float ---------------------
inline std::vector<float> blackbodyTable() {
  // Multi-lobe CIE 1931 observer fit, Wyman, Sloan & Shirley (JCGT 2013), Table 1.
  static const double lobes[3][3][4] = {
      {{0.362, 442, 0.0624, 0.0374}, {1.056, 599.8, 0.0264, 0.0323}, {-0.065, 501.1, 0.049, 0.0382}},
      {{0.821, 568.8, 0.0213, 0.0247}, {0.286, 530.9, 0.0613, 0.0322}, {0, 0, 0, 0}},
      {{1.217, 437, 0.0845, 0.0278}, {0.681, 459, 0.0385, 0.0725}, {0, 0, 0, 0}}};
  static const int lobeCount[3] = {3, 2, 2};
  double observer[95][3];
  for (int i = 0; i < 95; i++) {
    double wl = 360 + i * 5;
    for (int c = 0; c < 3; c++) {
      double sum = 0;
      for (int k = 0; k < lobeCount[c]; k++) {
        const double *l = lobes[c][k];
        double distance = (wl - l[1]) * (wl < l[1] ? l[2] : l[3]);
        sum = sum + l[0] * std::exp(-0.5 * distance * distance);
      }
      observer[i][c] = sum;
    }
  }
  auto planck = [](double nm, double kelvin) {
    double wl = nm * 1e-9;
    const double h = 6.62607015e-34, c = 299792458, k = 1.380649e-23;
    return (2 * h * c * c) / (std::pow(wl, 5) * std::expm1((h * c) / (wl * k * kelvin)));
  };
  auto rgbOf = [&](double kelvin, double rgb[3]) {
    double xyz[3] = {0, 0, 0};
    for (int i = 0; i < 95; i++) {
      double radiance = planck(360 + i * 5, kelvin) * 5e-9;
      for (int c = 0; c < 3; c++) xyz[c] += observer[i][c] * radiance;
    }
    rgb[0] = 3.2406 * xyz[0] - 1.5372 * xyz[1] - 0.4986 * xyz[2];
    rgb[1] = -0.9689 * xyz[0] + 1.8758 * xyz[1] + 0.0415 * xyz[2];
    rgb[2] = 0.0557 * xyz[0] - 0.204 * xyz[1] + 1.057 * xyz[2];
    for (int c = 0; c < 3; c++) rgb[c] = std::max(0.0, rgb[c]);
  };
  double ref[3];
  rgbOf(1800, ref);
  double referencePeak = std::max(ref[0], std::max(ref[1], ref[2]));
  std::vector<float> t(LUT_SIZE * 4);
  for (int i = 0; i < LUT_SIZE; i++) {
    double kelvin = 500 + ((10000 - 500) * (double)i) / (LUT_SIZE - 1);
    double rgb[3];
    rgbOf(kelvin, rgb);
    double peak = std::max(rgb[0], std::max(rgb[1], rgb[2]));
    float *o = &t[(size_t)i * 4];
    if (peak <= 0) {
      o[0] = o[1] = o[2] = 0;
      o[3] = -32;
      continue;
    }
    for (int c = 0; c < 3; c++) o[c] = (float)(rgb[c] / peak);
    o[3] = (float)std::max(-32.0, std::log2(peak / referencePeak));
  }
  return t;
}

// ---- parameters ---------------------------------------------------------------------
// RenderParams (volume-common.wgsl:4-10, packed at SceneVolumeRenderer.ts:329-345) and
// the camera uniforms (SceneVolumeRenderer.ts:440-456), as the kernels read them.
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
  bool halfRes;      // march at half resolution into depth slices first (VolumeSlices.ts)
};

inline RenderParams renderParams(double voxel, int B, const fp_camera &cam, const fp_render_settings &s,
                                 double time) {
  RenderParams p{};
  p.voxel = (float)voxel;
  int shift = 0;
  while ((1 << shift) < B) shift++;
  p.shift = shift;
  p.lightDivisor = s.lighting_divisor;
  int ls = 0;
  while ((1 << ls) < s.lighting_divisor) ls++;
  p.lightShift = shift - ls;
  p.brickSize = p.voxel * (float)B;
  p.texel = p.voxel * (float)s.lighting_divisor;
  p.flameOpacity = (float)s.flame_opacity;
  p.optical[0] = (float)s.smoke_density;
  p.optical[1] = (float)s.fire_intensity;
  p.optical[2] = (float)s.shadow_density;
  p.optical[3] = (float)s.scattering;
  for (int a = 0; a < 3; a++) {
    p.light[a] = (float)s.sun_direction[a];
    p.smoke[a] = (float)s.smoke_color[a];
    p.fire[a] = (float)s.fire_color[a];
    p.sceneLight[a] = (float)s.sun_color[a];
    p.ambient[a] = (float)s.ambient[a];
  }
  p.light[3] = (float)s.sun_intensity;
  p.sootGlow = (float)s.soot_glow;
  p.fire[3] = (float)s.temperature;
  p.filter = (float)s.filter;
  p.raySteps = (float)s.ray_steps;
  p.time = (float)time;
  for (int i = 0; i < 16; i++) p.ivp[i] = (float)cam.inverse_view_projection[i];
  p.width = cam.width;
  p.height = cam.height;
  p.halfRes = s.half_resolution != 0;
  return p;
}

// ---- boxes, page tables and pools --------------------------------------------------------
struct RBox {
  int origin[3], first, size[3];
};

struct Vec4 {
  float x, y, z, w;
};

// fireProGlow, fireProFlameAbsorption, fireProExtinction (fire-emission.wgsl:10-24).
inline float glowOf(float life) { return fmin(fmax(life * 5.0f, 0.0f), 1.0f); }
inline float extinctionOf(const Vec4 &f, const RenderParams &P) {
  return fmax(0.0f, f.x) * P.optical[0] + glowOf(f.z) * P.flameOpacity;
}

struct Renderer {
  const Sim<float> &sim;
  RenderParams P;
  int B, sF, side, nF, E, nR;   // field edge, shift, pool side, cells; render slot edge B+1, cells
  int lside, LE, nL;            // light brick edge, slot edge, texels
  int perSide, nO;              // occupancy blocks per axis and per slot
  // Caller-held state: the boxes and their page table, the blackbody table, and the
  // pools: render field (float4 x (B+1)^3 per slot), light (float4 x (B/ld+1)^3 per
  // slot), occupancy (one byte per 8^3 block: bit 0 flame, bit 1 visible).
  std::vector<RBox> &boxes;
  std::vector<int32_t> &table;
  const std::vector<float> &lut;
  std::vector<float> &R, &L;
  std::vector<uint8_t> &O;
  // Coarse smoke (smokeDivisor 2 or 4): the smoothed smoke with one-texel aprons,
  // (bs+1)^3 floats per slot, and the smoke lattice's brick edge, shift and cells.
  std::vector<float> &RS;
  bool coarse;
  int bs, sS, nS, ES, nRS;

  Renderer(const Sim<float> &s, const RenderParams &p, std::vector<RBox> &bx, std::vector<int32_t> &tb,
           const std::vector<float> &lt, std::vector<float> &r, std::vector<float> &l, std::vector<uint8_t> &o,
           std::vector<float> &rs)
      : sim(s), P(p), boxes(bx), table(tb), lut(lt), R(r), L(l), O(o), RS(rs) {
    coarse = s.smokeDiv > 1;
    bs = s.bs;
    sS = s.sS;
    nS = s.nS;
    ES = bs + 1;
    nRS = ES * ES * ES;
    B = s.B;
    sF = s.sF;
    side = s.side;
    nF = s.nF;
    E = B + 1;
    nR = E * E * E;
    lside = 1 << P.lightShift;
    LE = lside + 1;
    nL = LE * LE * LE;
    perSide = B / 8;
    nO = perSide * perSide * perSide;
  }

  // VolumeLighting.sync (VolumeLighting.ts:131-185): each box's first brick, first page
  // table entry and size; an empty page table.
  void syncBoxes() {
    boxes.clear();
    int first = 0;
    for (auto &bb : sim.renderBoxes) {
      RBox x;
      for (int a = 0; a < 3; a++) {
        x.origin[a] = bb.lo[a];
        x.size[a] = bb.hi[a] - bb.lo[a] + 1;
      }
      x.first = first;
      first += x.size[0] * x.size[1] * x.size[2];
      boxes.push_back(x);
    }
    table.assign((size_t)std::max(first, 1), -1);
  }

  // exportPages (lighting.wgsl:10-32): each brick id of a tile in a box writes its page.
  void exportPages() {
    int ids = sim.brickCount();
    for (int id = 0; id < ids; id++) {
      const int32_t *tile = &sim.renderTiles[(size_t)(id >> 6) * 4];
      if (tile[3] < 0) continue;
      const RBox &bx = boxes[tile[3]];
      int l = id & 63, loc[3] = {l & 3, (l >> 2) & 3, l >> 4}, local[3];
      bool inside = true;
      for (int a = 0; a < 3; a++) {
        local[a] = tile[a] * 4 + loc[a] - bx.origin[a];
        if (local[a] < 0 || local[a] >= bx.size[a]) inside = false;
      }
      if (!inside) continue;
      table[bx.first + (local[2] * bx.size[1] + local[1]) * bx.size[0] + local[0]] = sim.pages[id];
    }
  }

  // fireProPage (volume-bricks.wgsl:25-34).
  int pageAt(const RBox &bx, int x, int y, int z) const {
    int lx = x - bx.origin[0], ly = y - bx.origin[1], lz = z - bx.origin[2];
    if (lx < 0 || ly < 0 || lz < 0 || lx >= bx.size[0] || ly >= bx.size[1] || lz >= bx.size[2]) return -1;
    return table[bx.first + (lz * bx.size[1] + ly) * bx.size[0] + lx];
  }
  int slotOf(int page) const { return sim.pageSlot(page); }

  // The simulation's field cell through the box's page table (fireProLoad), zero without
  // a slot; and its channel x only.
  Vec4 fieldCell(const RBox &bx, int x, int y, int z) const {
    int pg = pageAt(bx, x >> sF, y >> sF, z >> sF);
    if (pg < 0) return {0, 0, 0, 0};
    int m = B - 1;
    const float *p = &sim.F[((size_t)slotOf(pg) * nF + (((z & m) * B + (y & m)) * B + (x & m))) * 4];
    return {p[0], p[1], p[2], p[3]};
  }

  // fireProField over the simulation's fields, channel x (lighting.wgsl:111-118 reads
  // the smoke channel): clamp to the box's outer cell centers, then fireProSample.
  float smokeSample(const RBox &bx, float px, float py, float pz) const {
    float sz = (float)(1 << sF);
    float q[3] = {px, py, pz};
    for (int a = 0; a < 3; a++) {
      float lo = (float)bx.origin[a] * sz + 0.5f;
      float hi = (float)(bx.origin[a] + bx.size[a]) * sz - 0.5f;
      q[a] = fmin(fmax(q[a], lo), hi);
    }
    int base[3];
    float f[3];
    for (int a = 0; a < 3; a++) {
      float g = std::floor(q[a] - 0.5f);
      base[a] = (int)g;
      f[a] = (q[a] - 0.5f) - g;
    }
    int last = B - 1;
    if (pageAt(bx, base[0] >> sF, base[1] >> sF, base[2] >> sF) < 0 && (base[0] & last) < last &&
        (base[1] & last) < last && (base[2] & last) < last)
      return 0.0f;
    float v = 0.0f;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          float wx = x ? f[0] : 1.0f - f[0], wy = y ? f[1] : 1.0f - f[1], wz = z ? f[2] : 1.0f - f[2];
          v += fieldCell(bx, base[0] + x, base[1] + y, base[2] + z).x * (wx * wy * wz);
        }
    return v;
  }

  // A texel of a pool with one-texel aprons: slot-major, (edge)^3 texels, x fastest.
  static const float *texel(const std::vector<float> &pool, int slot, int edge, int x, int y, int z) {
    return &pool[((size_t)slot * edge * edge * edge + ((size_t)z * edge + y) * edge + x) * 4];
  }

  // fireProField over the render field (volume-bricks.wgsl:91-104, :56-88): clamp, then a
  // trilinear fetch inside the base brick's slot (apron included) or, without a slot, zero
  // unless the sample reaches the next brick, where cells come from their own slots.
  Vec4 renderSample(const RBox &bx, float px, float py, float pz) const {
    float sz = (float)(1 << sF);
    float q[3] = {px, py, pz};
    for (int a = 0; a < 3; a++) {
      float lo = (float)bx.origin[a] * sz + 0.5f;
      float hi = (float)(bx.origin[a] + bx.size[a]) * sz - 0.5f;
      q[a] = fmin(fmax(q[a], lo), hi);
    }
    int base[3];
    float f[3];
    for (int a = 0; a < 3; a++) {
      float g = std::floor(q[a] - 0.5f);
      base[a] = (int)g;
      f[a] = (q[a] - 0.5f) - g;
    }
    int brick[3] = {base[0] >> sF, base[1] >> sF, base[2] >> sF};
    int pg = pageAt(bx, brick[0], brick[1], brick[2]);
    int last = B - 1;
    Vec4 v{0, 0, 0, 0};
    if (pg < 0 && (base[0] & last) < last && (base[1] & last) < last && (base[2] & last) < last) return v;
    int slot = pg >= 0 ? slotOf(pg) : -1;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          float wx = x ? f[0] : 1.0f - f[0], wy = y ? f[1] : 1.0f - f[1], wz = z ? f[2] : 1.0f - f[2];
          float w = wx * wy * wz;
          const float *t;
          if (slot >= 0) {
            t = texel(R, slot, E, base[0] + x - brick[0] * B, base[1] + y - brick[1] * B, base[2] + z - brick[2] * B);
          } else {
            int c[3] = {base[0] + x, base[1] + y, base[2] + z};
            int p2 = pageAt(bx, c[0] >> sF, c[1] >> sF, c[2] >> sF);
            if (p2 < 0) continue;
            t = texel(R, slotOf(p2), E, c[0] & last, c[1] & last, c[2] & last);
          }
          v.x += t[0] * w;
          v.y += t[1] * w;
          v.z += t[2] * w;
          v.w += t[3] * w;
        }
    return v;
  }

  // ---- coarse smoke ----
  // The simulation's coarse density at smoke cell (x, y, z) through the page table.
  float rawSmokeCell(const RBox &bx, int x, int y, int z) const {
    int pg = pageAt(bx, x >> sS, y >> sS, z >> sS);
    if (pg < 0) return 0.0f;
    int m = bs - 1;
    return sim.S[(size_t)slotOf(pg) * nS + (((z & m) * bs + (y & m)) * bs + (x & m))];
  }
  float &smokeTexel(int slot, int x, int y, int z) const {
    return RS[(size_t)slot * nRS + ((size_t)z * ES + y) * ES + x];
  }
  // fireProField on the smoke lattice (volume-bricks.wgsl:56-104), position in smoke cells,
  // over the simulation's density (`raw`) or the smoothed render smoke.
  float smokeField(const RBox &bx, float px, float py, float pz, bool raw) const {
    float sz = (float)bs;
    float q[3] = {px, py, pz};
    for (int a = 0; a < 3; a++)
      q[a] = fmin(fmax(q[a], (float)bx.origin[a] * sz + 0.5f), (float)(bx.origin[a] + bx.size[a]) * sz - 0.5f);
    int base[3], brick[3];
    float f[3];
    for (int a = 0; a < 3; a++) {
      float g = std::floor(q[a] - 0.5f);
      base[a] = (int)g;
      brick[a] = base[a] >> sS;
      f[a] = (q[a] - 0.5f) - g;
    }
    int pg = pageAt(bx, brick[0], brick[1], brick[2]);
    int last = bs - 1;
    if (pg < 0 && (base[0] & last) < last && (base[1] & last) < last && (base[2] & last) < last) return 0.0f;
    int slot = pg >= 0 ? slotOf(pg) : -1;
    float v = 0.0f;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          float w = (x ? f[0] : 1.0f - f[0]) * (y ? f[1] : 1.0f - f[1]) * (z ? f[2] : 1.0f - f[2]);
          int c[3] = {base[0] + x, base[1] + y, base[2] + z};
          float t;
          if (raw) {
            t = rawSmokeCell(bx, c[0], c[1], c[2]);
          } else if (slot >= 0) {
            t = smokeTexel(slot, c[0] - brick[0] * bs, c[1] - brick[1] * bs, c[2] - brick[2] * bs);
          } else {
            int p2 = pageAt(bx, c[0] >> sS, c[1] >> sS, c[2] >> sS);
            if (p2 < 0) continue;
            t = smokeTexel(slotOf(p2), c[0] & last, c[1] & last, c[2] & last);
          }
          v += t * w;
        }
    return v;
  }
  // renderSmokeAt (volume-common.wgsl:74-78): the smoothed smoke at a field position.
  float renderSmokeAt(const RBox &bx, float fx, float fy, float fz) const {
    float ratio = (float)(B / bs);
    return smokeField(bx, fx / ratio, fy / ratio, fz / ratio, false);
  }
  // cachedSmokeCell (lighting.wgsl:133-152).
  float cachedSmokeCell(const RBox &bx, int x, int y, int z) const {
    int m = bs - 1;
    int b[3] = {x >> sS, y >> sS, z >> sS}, l[3] = {x & m, y & m, z & m};
    int pg = pageAt(bx, b[0], b[1], b[2]);
    if (pg >= 0) return smokeTexel(slotOf(pg), l[0], l[1], l[2]);
    for (int mk = 1; mk < 8; mk++) {
      int st[3] = {mk & 1, (mk >> 1) & 1, (mk >> 2) & 1};
      bool skip = false;
      for (int a = 0; a < 3; a++) skip = skip || (st[a] && l[a] != 0);
      if (skip) continue;
      int prev = pageAt(bx, b[0] - st[0], b[1] - st[1], b[2] - st[2]);
      if (prev >= 0) return smokeTexel(slotOf(prev), l[0] + st[0] * bs, l[1] + st[1] * bs, l[2] + st[2] * bs);
    }
    return 0.0f;
  }

  // sampleWorld (volume-common.wgsl:41-51): zero outside the box, else the render field
  // at p in cells, its smoke from the coarse grid when smoke has one.
  Vec4 sampleWorld(const RBox &bx, float px, float py, float pz) const {
    float size = P.brickSize;
    float p[3] = {px, py, pz};
    for (int a = 0; a < 3; a++)
      if (p[a] < (float)bx.origin[a] * size || p[a] > (float)(bx.origin[a] + bx.size[a]) * size)
        return {0, 0, 0, 0};
    float fx = px / P.voxel, fy = py / P.voxel, fz = pz / P.voxel;
    Vec4 f = renderSample(bx, fx, fy, fz);
    if (coarse) f.x = renderSmokeAt(bx, fx, fy, fz);
    return f;
  }

  // The blackbody texel pair at kelvin, linearly filtered (fire-emission.wgsl:43-45).
  void spectrum(float kelvin, float out[4]) const {
    float u = fmin(fmax((kelvin - 500.0f) / 9500.0f, 0.0f), 1.0f) * (float)(LUT_SIZE - 1);
    float i0f = std::floor(u);
    float fr = u - i0f;
    int i0 = (int)i0f, i1 = std::min(i0 + 1, LUT_SIZE - 1);
    for (int c = 0; c < 4; c++) out[c] = lut[(size_t)i0 * 4 + c] * (1.0f - fr) + lut[(size_t)i1 * 4 + c] * fr;
  }

  // fireProEmission (fire-emission.wgsl:22-56) with smoke = (albedo, soot glow).
  void emission(const Vec4 &f, float out[3]) const {
    out[0] = out[1] = out[2] = 0.0f;
    if (P.optical[1] <= 0.0f) return;
    float dens = fmax(0.0f, f.x);
    float flame = glowOf(f.z) * P.flameOpacity;
    float source[3];
    for (int c = 0; c < 3; c++) source[c] = ((1.0f - P.smoke[c]) * dens) * P.optical[0] * P.sootGlow + flame;
    if (source[0] == 0.0f && source[1] == 0.0f && source[2] == 0.0f) return;
    float heat = fmax(0.0f, f.y);
    float kelvin = 800.0f + heat / (heat + 2.0f) * (P.fire[3] - 800.0f);
    float s[4];
    spectrum(kelvin, s);
    float scale = fexp2(fmin(fmax(s[3], -24.0f), 10.0f));
    float thermal[3] = {s[0] * scale, s[1] * scale, s[2] * scale};
    float tintPeak = fmax(P.fire[0], fmax(P.fire[1], P.fire[2]));
    float saturation = 1.0f - fmin(P.fire[0], fmin(P.fire[1], P.fire[2])) / fmax(tintPeak, 0.000001f);
    float peak = fmax(thermal[0], fmax(thermal[1], thermal[2]));
    for (int c = 0; c < 3; c++) {
      float colored = mixv(thermal[c] * tintPeak, P.fire[c] * peak, saturation);
      out[c] = colored * source[c] * P.optical[1];
    }
  }

  // exitDistance (lighting.wgsl:258-266).
  float exitDistance(const RBox &bx, const float p[3], const float d[3]) const {
    float best = 1e30f;
    float dist[3];
    for (int a = 0; a < 3; a++) {
      float edge = (float)(d[a] > 0.0f ? bx.origin[a] + bx.size[a] : bx.origin[a]) * P.brickSize;
      bool moving = fabs_(d[a]) > 0.000001f;
      dist[a] = moving ? (edge - p[a]) / d[a] : 1e30f;
    }
    best = fmin(dist[0], fmin(dist[1], dist[2]));
    return fmax(0.0f, best);
  }

  // incidentLight (lighting.wgsl:251-300).
  Vec4 incidentLight(const RBox &bx, const float p[3]) const {
    Vec4 lf = sampleWorld(bx, p[0], p[1], p[2]);
    if (lf.x + glowOf(lf.z) < 0.00001f) return {0, 0, 0, 1};
    float sun[3];
    normalize3(P.light, sun);
    float ds = exitDistance(bx, p, sun) / (float)SHADOW_STEPS;
    float depth = 0.0f;
    if (P.light[3] > 0.0f) {
      for (int i = 0; i < SHADOW_STEPS; i++) {
        float t = ((float)i + 0.5f) * ds;
        Vec4 f = sampleWorld(bx, p[0] + sun[0] * t, p[1] + sun[1] * t, p[2] + sun[2] * t);
        depth += (fmax(0.0f, f.x) * P.optical[2] + glowOf(f.z) * P.flameOpacity) * ds;
        if (depth > 9.0f) break;
      }
    }
    static const float dirs[6][3] = {{.36f, .48f, -.8f}, {-.36f, -.48f, .8f}, {-.8f, .6f, 0.0f},
                                     {.8f, -.6f, 0.0f},  {.48f, .64f, .6f},   {-.48f, -.64f, -.6f}};
    float glow[3] = {0, 0, 0};
    if (P.optical[1] > 0.0f && lf.x > 0.00001f) {
      for (int d = 0; d < 6; d++) {
        float t = 0.0f, interval = 0.075f, transmission = 1.0f;
        for (int s = 0; s < 6; s++) {
          float at = t + interval * 0.5f;
          Vec4 f = sampleWorld(bx, p[0] + dirs[d][0] * at, p[1] + dirs[d][1] * at, p[2] + dirs[d][2] * at);
          float ext = extinctionOf(f, P);
          float att = fexp(-ext * interval);
          float e[3];
          emission(f, e);
          float it = integratedTransmission(ext, interval, att);
          for (int c = 0; c < 3; c++) glow[c] += transmission * e[c] * it / 6.0f;
          transmission *= att;
          t += interval;
          interval *= 1.65f;
        }
      }
    }
    return {glow[0], glow[1], glow[2], fexp(-depth)};
  }

  // integratedTransmission (volume-common.wgsl:55-61).
  static float integratedTransmission(float ext, float distance, float att) {
    float depth = ext * distance;
    if (depth < 0.001f) return distance * (1.0f - 0.5f * depth + depth * depth / 6.0f);
    return (1.0f - att) / ext;
  }

  static void normalize3(const float v[3], float out[3]) {
    float len = length3(v[0], v[1], v[2]);
    for (int a = 0; a < 3; a++) out[a] = v[a] / len;
  }

  // A brick of the active list that has a slot inside a box (lighting.wgsl:34-60).
  bool slotted(int id, int &page, int &boxIndex, int brick[3]) const {
    page = sim.pages[id];
    const int32_t *tile = &sim.renderTiles[(size_t)(id >> 6) * 4];
    if (page < 0 || tile[3] < 0) return false;
    boxIndex = tile[3];
    const RBox &bx = boxes[boxIndex];
    int l = id & 63, loc[3] = {l & 3, (l >> 2) & 3, l >> 4};
    for (int a = 0; a < 3; a++) {
      brick[a] = tile[a] * 4 + loc[a];
      int local = brick[a] - bx.origin[a];
      if (local < 0 || local >= bx.size[a]) return false;
    }
    return true;
  }

  // smoothSmoke (lighting.wgsl:62-97): every texel of the slot and its apron.
  void smoothBrick(int page, const RBox &bx, const int brick[3]) {
    int slot = slotOf(page);
    for (int z = 0; z < E; z++)
      for (int y = 0; y < E; y++)
        for (int x = 0; x < E; x++) {
          int cell[3] = {brick[0] * B + x, brick[1] * B + y, brick[2] * B + z};
          Vec4 original = fieldCell(bx, cell[0], cell[1], cell[2]);
          if (coarse) {
            // The fields as they are (VolumeLighting.ts:206-208).
            float *o = const_cast<float *>(texel(R, slot, E, x, y, z));
            o[0] = original.x;
            o[1] = original.y;
            o[2] = original.z;
            o[3] = original.w;
            continue;
          }
          float density = 0.0f;
          for (int corner = 0; corner < 8; corner++)
            density += smokeSample(bx, (float)cell[0] + (float)(corner & 1), (float)cell[1] + (float)((corner >> 1) & 1),
                                   (float)cell[2] + (float)((corner >> 2) & 1));
          float *o = const_cast<float *>(texel(R, slot, E, x, y, z));
          o[0] = fmax(0.0f, density * 0.125f);
          o[1] = original.y;
          o[2] = original.z;
          o[3] = original.w;
        }
  }

  // smoothCoarseSmoke (lighting.wgsl:105-129): every texel of the slot and its apron.
  void smoothCoarseBrick(int page, const RBox &bx, const int brick[3]) {
    int slot = slotOf(page);
    for (int z = 0; z < ES; z++)
      for (int y = 0; y < ES; y++)
        for (int x = 0; x < ES; x++) {
          int cell[3] = {brick[0] * bs + x, brick[1] * bs + y, brick[2] * bs + z};
          float density = 0.0f;
          for (int corner = 0; corner < 8; corner++)
            density += smokeField(bx, (float)cell[0] + (float)(corner & 1), (float)cell[1] + (float)((corner >> 1) & 1),
                                  (float)cell[2] + (float)((corner >> 2) & 1), true);
          smokeTexel(slot, x, y, z) = fmax(0.0f, density * 0.125f);
        }
  }

  // buildOccupancy (lighting.wgsl:150-253): per 8^3 block, the 12^3 cells of the block and
  // its two-cell halo whose bricks have slots. With coarse smoke the fine cells give flame
  // only, and the block is visible when its largest glow plus the largest smoothed smoke
  // among its smoke donors (clamped to the box) passes the cutoff.
  void occupancyBrick(int page, const RBox &bx, const int brick[3]) {
    int slot = slotOf(page);
    for (int block = 0; block < nO; block++) {
      int b[3] = {block % perSide, (block / perSide) % perSide, block / (perSide * perSide)};
      uint8_t flags = 0;
      float maxGlow = 0.0f;
      for (int unit = 0; unit < 216; unit++) {
        int cell[3] = {brick[0] * B + b[0] * 8 - 2 + (unit % 6) * 2, brick[1] * B + b[1] * 8 - 2 + ((unit / 6) % 6) * 2,
                       brick[2] * B + b[2] * 8 - 2 + (unit / 36) * 2};
        int pg = pageAt(bx, cell[0] >> sF, cell[1] >> sF, cell[2] >> sF);
        if (pg < 0) continue;
        int m = B - 1, s2 = slotOf(pg);
        for (int corner = 0; corner < 8; corner++) {
          const float *f = texel(R, s2, E, (cell[0] & m) + (corner & 1), (cell[1] & m) + ((corner >> 1) & 1),
                                 (cell[2] & m) + ((corner >> 2) & 1));
          float glow = glowOf(f[2]);
          float smoke = coarse ? 0.0f : f[0];
          if (glow > 0.002f) flags |= 1;
          if (smoke + glow > 0.002f) flags |= 2;
          maxGlow = fmax(maxGlow, glow);
        }
      }
      if (coarse) {
        float ratio = (float)(B / bs), maxSmoke = 0.0f;
        int lo[3], hi[3], boxLo[3], boxHi[3];
        for (int a = 0; a < 3; a++) {
          int origin = brick[a] * B + b[a] * 8 - 2;
          lo[a] = (int)std::floor((float)origin / ratio - 0.5f);
          hi[a] = (int)std::floor((float)(origin + 12) / ratio - 0.5f) + 1;
          boxLo[a] = bx.origin[a] * bs;
          boxHi[a] = (bx.origin[a] + bx.size[a]) * bs - 1;
        }
        for (int z = lo[2]; z <= hi[2]; z++)
          for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++) {
              int c[3] = {x, y, z};
              for (int a = 0; a < 3; a++) c[a] = std::min(std::max(c[a], boxLo[a]), boxHi[a]);
              maxSmoke = fmax(maxSmoke, cachedSmokeCell(bx, c[0], c[1], c[2]));
            }
        if (maxSmoke + maxGlow > 0.002f) flags |= 2;
      }
      O[(size_t)slot * nO + block] = flags;
    }
  }

  // lightSlot (lighting.wgsl:302-345): the incident light at every texel of the slot and
  // its apron, at texel centers of the lighting lattice.
  void lightBrick(int page, const RBox &bx, const int brick[3]) {
    int slot = slotOf(page);
    for (int z = 0; z < LE; z++)
      for (int y = 0; y < LE; y++)
        for (int x = 0; x < LE; x++) {
          float p[3] = {((float)(brick[0] * lside + x) + 0.5f) * P.texel, ((float)(brick[1] * lside + y) + 0.5f) * P.texel,
                        ((float)(brick[2] * lside + z) + 0.5f) * P.texel};
          Vec4 v = incidentLight(bx, p);
          float *o = const_cast<float *>(texel(L, slot, LE, x, y, z));
          o[0] = v.x;
          o[1] = v.y;
          o[2] = v.z;
          o[3] = v.w;
        }
  }

  enum LightWork { LIGHT_NONE = 0, LIGHT_ALL = 1, LIGHT_FRESH = 2 };

  // The kernels of one VolumeLighting.encode (VolumeLighting.ts:339-367), in order: page
  // export, smoothing and occupancy of every listed brick, then the light of every listed
  // brick or of the slots given out this step.
  void prepare(bool pages, bool smooth, bool occupancy, int lightWork) {
    if (pages) exportPages();
    if (smooth)
      for (int id : sim.activeList) {
        int page, bi, brick[3];
        if (!slotted(id, page, bi, brick)) continue;
        smoothBrick(page, boxes[bi], brick);
        if (coarse) smoothCoarseBrick(page, boxes[bi], brick);
      }
    if (occupancy)
      for (int id : sim.activeList) {
        int page, bi, brick[3];
        if (!slotted(id, page, bi, brick)) continue;
        occupancyBrick(page, boxes[bi], brick);
      }
    if (lightWork == LIGHT_ALL)
      for (int id : sim.activeList) {
        int page, bi, brick[3];
        if (!slotted(id, page, bi, brick)) continue;
        lightBrick(page, boxes[bi], brick);
      }
    if (lightWork == LIGHT_FRESH)
      for (auto &z : sim.zeroList) {
        if (!(z.second & 0x80000000u)) continue;
        int page, bi, brick[3];
        if (!slotted(z.first, page, bi, brick)) continue;
        lightBrick(page, boxes[bi], brick);
      }
  }

  // ---- the camera march ----------------------------------------------------------------
  // fireProHit (volume-march.wgsl:27-35).
  void hit(const float o[3], const float d[3], const RBox &bx, float &entry, float &exit) const {
    float lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
      float safe = fabs_(d[a]) > 0.000001f ? d[a] : 0.000001f;
      float u = ((float)bx.origin[a] * P.brickSize - o[a]) / safe;
      float v = ((float)(bx.origin[a] + bx.size[a]) * P.brickSize - o[a]) / safe;
      lo[a] = fmin(u, v);
      hi[a] = fmax(u, v);
    }
    entry = fmax(0.0f, fmax(lo[0], fmax(lo[1], lo[2])));
    exit = fmin(hi[0], fmin(hi[1], hi[2]));
  }

  // fireProLight (volume-bricks.wgsl:160-179).
  Vec4 lightSample(const RBox &bx, float px, float py, float pz) const {
    float sz = (float)lside;
    float q[3] = {px, py, pz};
    for (int a = 0; a < 3; a++) {
      float lo = (float)bx.origin[a] * sz + 0.5f;
      q[a] = fmin(fmax(q[a], lo), (float)(bx.origin[a] + bx.size[a]) * sz - 0.5f);
    }
    int base[3], brick[3];
    float f[3];
    for (int a = 0; a < 3; a++) {
      float g = std::floor(q[a] - 0.5f);
      base[a] = (int)g;
      brick[a] = base[a] >> P.lightShift;
      f[a] = (q[a] - 0.5f) - g;
    }
    int pg = pageAt(bx, brick[0], brick[1], brick[2]);
    if (pg < 0) return {0, 0, 0, 1};
    int slot = slotOf(pg);
    Vec4 v{0, 0, 0, 0};
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          float wx = x ? f[0] : 1.0f - f[0], wy = y ? f[1] : 1.0f - f[1], wz = z ? f[2] : 1.0f - f[2];
          float w = wx * wy * wz;
          const float *t = texel(L, slot, LE, base[0] + x - brick[0] * lside, base[1] + y - brick[1] * lside,
                                 base[2] + z - brick[2] * lside);
          v.x += t[0] * w;
          v.y += t[1] * w;
          v.z += t[2] * w;
          v.w += t[3] * w;
        }
    return v;
  }

  // fireProReconstruct (volume-reconstruction.wgsl:5-96).
  Vec4 reconstruct(const RBox &bx, const float c[3], float mode) const {
    if (mode < 0.5f) return renderSample(bx, c[0], c[1], c[2]);
    float pos[2][3], wt[2][3];
    for (int a = 0; a < 3; a++) {
      float grid = c[a] - 0.5f;
      float base = std::floor(grid);
      float q = grid - base;
      if (mode < 1.5f) {
        float center = std::floor(grid + 0.5f);
        float delta = grid - center;
        float w0 = 0.5f * (0.5f - delta) * (0.5f - delta);
        float w1 = 0.75f - delta * delta;
        float w2 = 0.5f * (0.5f + delta) * (0.5f + delta);
        wt[0][a] = w0 + w1;
        wt[1][a] = w2;
        pos[0][a] = center - 0.5f + w1 / wt[0][a];
        pos[1][a] = center + 1.5f;
      } else {
        float inv = 1.0f - q;
        float w0 = inv * inv * inv / 6.0f;
        float w1 = (3.0f * q * q * q - 6.0f * q * q + 4.0f) / 6.0f;
        float w2 = (-3.0f * q * q * q + 3.0f * q * q + 3.0f * q + 1.0f) / 6.0f;
        float w3 = q * q * q / 6.0f;
        wt[0][a] = w0 + w1;
        wt[1][a] = w2 + w3;
        pos[0][a] = base - 0.5f + w1 / wt[0][a];
        pos[1][a] = base + 1.5f + w3 / wt[1][a];
      }
    }
    Vec4 v{0, 0, 0, 0};
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        for (int k = 0; k < 2; k++) {
          float w = wt[i][0] * wt[j][1] * wt[k][2];
          Vec4 s = renderSample(bx, pos[i][0], pos[j][1], pos[k][2]);
          v.x += w * s.x;
          v.y += w * s.y;
          v.z += w * s.z;
          v.w += w * s.w;
        }
    return {fmax(v.x, 0.0f), fmax(v.y, 0.0f), fmax(v.z, 0.0f), fmax(v.w, 0.0f)};
  }

  // fireProJitter (volume-march.wgsl:92-100).
  static float jitter(uint32_t px, uint32_t py, float time) {
    uint32_t s[3] = {px, py, (uint32_t)(time * 60.0f)};
    for (int a = 0; a < 3; a++) s[a] = s[a] * 1664525u + 1013904223u;
    s[0] += s[1] * s[2];
    s[1] += s[2] * s[0];
    s[2] += s[0] * s[1];
    for (int a = 0; a < 3; a++) s[a] ^= s[a] >> 16u;
    s[0] += s[1] * s[2];
    return (float)(s[0] >> 8u) / 16777216.0f;
  }

  // M * (x, y, z, w) with M column-major, summed column by column.
  void transform(const float v[4], float out[4]) const {
    for (int r = 0; r < 4; r++)
      out[r] = ((P.ivp[r] * v[0] + P.ivp[4 + r] * v[1]) + P.ivp[8 + r] * v[2]) + P.ivp[12 + r] * v[3];
  }

  // fireProSceneVolume's beauty path for pixel (px, py), row 0 at the top.
  // fireProRay (volume-march.wgsl:18-24) through screen position uv.
  void ray(const float uv[2], float o[3], float d[3]) const {
    float clip[2] = {uv[0] * 2.0f + -1.0f, uv[1] * -2.0f + 1.0f};
    float nearIn[4] = {clip[0], clip[1], 0.0f, 1.0f}, farIn[4] = {clip[0], clip[1], 0.99999f, 1.0f};
    float nh[4], fh[4];
    transform(nearIn, nh);
    transform(farIn, fh);
    float dd[3];
    for (int a = 0; a < 3; a++) o[a] = nh[a] / nh[3];
    for (int a = 0; a < 3; a++) dd[a] = fh[a] / fh[3] - o[a];
    normalize3(dd, d);
  }
  // fireProSpan (volume-march.wgsl:39-48).
  void span(const float o[3], const float d[3], float &sx, float &sy, float &sz) const {
    sx = 1e30f;
    sy = 0.0f;
    sz = 0.0f;
    for (const RBox &bx : boxes) {
      float en, ex;
      hit(o, d, bx, en, ex);
      if (ex > en) {
        sx = fmin(sx, en);
        sy = fmax(sy, ex);
        sz = sz + ex - en;
      }
    }
  }
  // fireProDirect (volume-march.wgsl:138-143).
  void directLight(const float d[3], float direct[3]) const {
    float ln[3];
    normalize3(P.light, ln);
    float g = P.optical[3];
    float mu = d[0] * ln[0] + d[1] * ln[1] + d[2] * ln[2];
    float base = fmax(0.05f, 1.0f + g * g - 2.0f * g * mu);
    float phase = (1.0f - g * g) / (base * std::sqrt(base));
    for (int a = 0; a < 3; a++) direct[a] = P.sceneLight[a] * P.light[3] * (0.35f + phase * 0.3f);
  }

  // The emission and extinction of a sample (fireProShade, volume-march.wgsl:103-138)
  // integrated over ds into the ray's state.
  void shade(const RBox &bx, const float cell[3], uint8_t flags, const float direct[3], float ds,
             float &transmittance, float radiance[3]) const {
    Vec4 f = reconstruct(bx, cell, (flags & 1) ? P.filter : 0.0f);
    // fireProRenderField (volume-reconstruction.wgsl:100-116): smoke from its own lattice.
    if (coarse) f.x = renderSmokeAt(bx, cell[0], cell[1], cell[2]);
    if (!(f.x + glowOf(f.z) > 0.002f)) return;
    float lightRatio = (float)P.lightDivisor;
    float glow[3];
    emission(f, glow);
    float density = fmax(0.0f, f.x) * P.optical[0];
    float scat[3] = {0, 0, 0};
    if (density > 0.0f) {
      Vec4 inc = lightSample(bx, cell[0] / lightRatio, cell[1] / lightRatio, cell[2] / lightRatio);
      float irgb[3] = {inc.x, inc.y, inc.z};
      for (int a = 0; a < 3; a++) scat[a] = P.smoke[a] * (P.ambient[a] + direct[a] * inc.w + irgb[a] * 0.65f) * density;
    }
    float ext = extinctionOf(f, P);
    float tau = ext * ds;
    float integral = ds * (1.0f - 0.5f * tau + tau * tau / 6.0f);
    if (tau >= 0.001f) integral = (1.0f - fexp(-tau)) / ext;
    for (int a = 0; a < 3; a++) radiance[a] += transmittance * (scat[a] + glow[a]) * integral;
    transmittance *= 1.0f - fmin(1.0f, ext * integral);
  }
  // The 8-cell block flags of the sample at p (cell c) in box bx and the skip size.
  uint8_t blockFlags(const RBox &bx, const int c[3], float &size) const {
    int pg = pageAt(bx, c[0] >> P.shift, c[1] >> P.shift, c[2] >> P.shift);
    size = P.brickSize;
    if (pg < 0) return 0;
    int m = B - 1;
    int blk = (((c[2] & m) >> 3) * perSide + ((c[1] & m) >> 3)) * perSide + ((c[0] & m) >> 3);
    size = P.voxel * 8.0f;
    return O[(size_t)slotOf(pg) * nO + blk];
  }

  static constexpr int SLICES = 8;  // VolumeSlices.ts DEPTH_SLICES
  // buildSlices (volume-slices.wgsl:27-122) for texel (px, py) of an sw x sh grid: the
  // march's state at SLICES evenly spaced depths from the first entry to the last exit,
  // into out[k * sw * sh].
  // `visits` counts the sample positions looked up in the occupancy, `shades` those
  // reconstructed and shaded (the render's sample counts).
  void slice(int px, int py, int sw, int sh, float *out, uint64_t &visits, uint64_t &shades) const {
    float uv[2] = {((float)px + 0.5f) / (float)sw, ((float)py + 0.5f) / (float)sh};
    float o[3], d[3], sx, sy, sz;
    ray(uv, o, d);
    span(o, d, sx, sy, sz);
    float spacing = fmax(sy - sx, 0.0f) / (float)SLICES;
    float stepSize = fmax(P.voxel * 0.5f, sz / P.raySteps);
    float jit = jitter((uint32_t)px, (uint32_t)py, P.time);
    float direct[3];
    directLight(d, direct);
    float safe[3], exitSide[3];
    for (int a = 0; a < 3; a++) {
      safe[a] = fabs_(d[a]) > 0.000001f ? d[a] : 0.000001f;
      exitSide[a] = d[a] > 0.0f ? 1.0f : 0.0f;
    }
    float t = sx, transmittance = 1.0f, radiance[3] = {0, 0, 0};
    int recorded = 0, cur = -1;
    float depth = sx + spacing, curExit = t;
    auto store = [&](int k) {
      float *v = out + (size_t)k * sw * sh * 4;
      v[0] = radiance[0];
      v[1] = radiance[1];
      v[2] = radiance[2];
      v[3] = 1.0f - transmittance;
    };
    while (!(recorded >= SLICES || t >= sy || transmittance < 0.008f)) {
      if (t >= curExit) {
        float nx = 1e30f, ny = 0.0f;
        int ni = -1;
        for (int i = 0; i < (int)boxes.size(); i++) {
          float en, ex;
          hit(o, d, boxes[i], en, ex);
          if (ex > en && ex > t && en < nx) {
            nx = en;
            ny = ex;
            ni = i;
          }
        }
        if (ni < 0) break;
        cur = ni;
        curExit = ny;
        t = fmax(t, nx);
      } else {
        const RBox &bx = boxes[cur];
        visits++;
        // Steps stop at the next recorded depth and at the box's exit.
        float ds = fmin(stepSize, fmin(depth, curExit) - t);
        float at = t + jit * ds;
        float p[3] = {o[0] + d[0] * at, o[1] + d[1] * at, o[2] + d[2] * at};
        float cell[3] = {p[0] / P.voxel, p[1] / P.voxel, p[2] / P.voxel};
        int c[3] = {(int)std::floor(cell[0]), (int)std::floor(cell[1]), (int)std::floor(cell[2])};
        float size;
        uint8_t flags = blockFlags(bx, c, size);
        if (!(flags & 2)) {
          float dist[3];
          for (int a = 0; a < 3; a++) dist[a] = ((std::floor(p[a] / size) + exitSide[a]) * size - p[a]) / safe[a];
          float distance = fmin(dist[0], fmin(dist[1], dist[2]));
          t = fmin(t + fmax(1.0f, std::floor(distance / stepSize) + 1.0f) * stepSize, curExit);
        } else {
          shades++;
          shade(bx, cell, flags, direct, ds, transmittance, radiance);
          t += ds;
        }
      }
      while (recorded < SLICES && !(t < depth)) {
        store(recorded);
        recorded++;
        depth = sx + spacing * (float)(recorded + 1);
      }
    }
    for (int k = recorded; k < SLICES; k++) store(k);
  }

  void pixel(int px, int py, float out[4], const float *slices = nullptr, int sw = 0, int sh = 0) const {
    uint64_t visits = 0, shades = 0;
    pixel(px, py, out, slices, sw, sh, visits, shades);
  }
  void pixel(int px, int py, float out[4], const float *slices, int sw, int sh, uint64_t &visits,
             uint64_t &shades) const {
    out[0] = out[1] = out[2] = out[3] = 0.0f;
    float uv[2] = {((float)px + 0.5f) / (float)P.width, ((float)py + 0.5f) / (float)P.height};
    float clip[2] = {uv[0] * 2.0f + -1.0f, uv[1] * -2.0f + 1.0f};
    // fireProRay (volume-march.wgsl:18-24)
    float nearIn[4] = {clip[0], clip[1], 0.0f, 1.0f}, farIn[4] = {clip[0], clip[1], 0.99999f, 1.0f};
    float nh[4], fh[4];
    transform(nearIn, nh);
    transform(farIn, fh);
    float o[3], d[3], dd[3];
    for (int a = 0; a < 3; a++) o[a] = nh[a] / nh[3];
    for (int a = 0; a < 3; a++) dd[a] = fh[a] / fh[3] - o[a];
    normalize3(dd, d);
    // fireProSpan (volume-march.wgsl:39-48)
    float sx = 1e30f, sy = 0.0f, sz = 0.0f;
    for (const RBox &bx : boxes) {
      float en, ex;
      hit(o, d, bx, en, ex);
      if (ex > en) {
        sx = fmin(sx, en);
        sy = fmax(sy, ex);
        sz = sz + ex - en;
      }
    }
    if (sx >= sy) return;
    float end = sy;
    if (end <= sx) return;
    float stepSize = fmax(P.voxel * 0.5f, sz / P.raySteps);
    float jit = jitter((uint32_t)px, (uint32_t)py, P.time);
    float direct[3];
    directLight(d, direct);
    float t0 = sx, trans0 = 1.0f, rad0[3] = {0, 0, 0};
    if (slices) {
      // scene-volume.wgsl:147-161: resume from the last stored depth in front of the scene
      // (with no scene depth, the last one), a linear fetch at (uv, (k - 1/2) / count):
      // one layer, bilinear in x and y, clamped to the edge.
      float count = (float)SLICES;
      float s = std::floor((end - sx) / (sy - sx) * count);
      if (s >= 1.0f) {
        float k = fmin(s, count);
        float fx0 = uv[0] * (float)sw - 0.5f, fy0 = uv[1] * (float)sh - 0.5f;
        float gx = std::floor(fx0), gy = std::floor(fy0);
        float fx = fx0 - gx, fy = fy0 - gy;
        int x0 = std::min(std::max((int)gx, 0), sw - 1), x1 = std::min(std::max((int)gx + 1, 0), sw - 1);
        int y0 = std::min(std::max((int)gy, 0), sh - 1), y1 = std::min(std::max((int)gy + 1, 0), sh - 1);
        const float *layer = slices + (size_t)((int)k - 1) * sw * sh * 4;
        auto at = [&](int x, int y, int ch) { return layer[((size_t)y * sw + x) * 4 + ch]; };
        float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
        float v[4];
        for (int ch = 0; ch < 4; ch++)
          v[ch] = ((at(x0, y0, ch) * w00 + at(x1, y0, ch) * w10) + at(x0, y1, ch) * w01) + at(x1, y1, ch) * w11;
        t0 = sx + (sy - sx) * k / count;
        trans0 = 1.0f - v[3];
        for (int a = 0; a < 3; a++) rad0[a] = v[a];
      }
    }
    march(o, d, t0, trans0, rad0, end, stepSize, jit, direct, out, visits, shades);
  }

  // fireProMarch (volume-march.wgsl:183-266) from t0 with transmittance trans0 and
  // radiance rad0.
  void march(const float o[3], const float d[3], float t0, float trans0, const float rad0[3], float end,
             float stepSize, float jit, const float direct[3], float out[4], uint64_t &visits,
             uint64_t &shades) const {
    float safe[3], exitSide[3];
    for (int a = 0; a < 3; a++) {
      safe[a] = fabs_(d[a]) > 0.000001f ? d[a] : 0.000001f;
      exitSide[a] = d[a] > 0.0f ? 1.0f : 0.0f;
    }
    float t = t0, transmittance = trans0, radiance[3] = {rad0[0], rad0[1], rad0[2]};
    int cur = -1;
    float curExit = t;
    float lightRatio = (float)P.lightDivisor;
    while (true) {
      if (t >= end || transmittance < 0.008f) break;
      if (t >= curExit) {
        // fireProEnter / fireProNext (volume-march.wgsl:53-68, :159-174)
        float nx = 1e30f, ny = 0.0f;
        int ni = -1;
        for (int i = 0; i < (int)boxes.size(); i++) {
          float en, ex;
          hit(o, d, boxes[i], en, ex);
          if (ex > en && ex > t && en < nx) {
            nx = en;
            ny = ex;
            ni = i;
          }
        }
        if (ni < 0) break;
        cur = ni;
        curExit = ny;
        t = fmax(t, nx);
        continue;
      }
      const RBox &bx = boxes[cur];
      visits++;
      float ds = fmin(stepSize, fmin(end, curExit) - t);
      float at = t + jit * ds;
      float p[3] = {o[0] + d[0] * at, o[1] + d[1] * at, o[2] + d[2] * at};
      float cell[3] = {p[0] / P.voxel, p[1] / P.voxel, p[2] / P.voxel};
      int c[3] = {(int)std::floor(cell[0]), (int)std::floor(cell[1]), (int)std::floor(cell[2])};
      int pg = pageAt(bx, c[0] >> P.shift, c[1] >> P.shift, c[2] >> P.shift);
      float size = P.brickSize;
      uint8_t flags = 0;
      if (pg >= 0) {
        int m = B - 1;
        int blk = (((c[2] & m) >> 3) * perSide + ((c[1] & m) >> 3)) * perSide + ((c[0] & m) >> 3);
        flags = O[(size_t)slotOf(pg) * nO + blk];
        size = P.voxel * 8.0f;
      }
      if (!(flags & 2)) {
        float dist[3];
        for (int a = 0; a < 3; a++) {
          float edge = (std::floor(p[a] / size) + exitSide[a]) * size;
          dist[a] = (edge - p[a]) / safe[a];
        }
        float distance = fmin(dist[0], fmin(dist[1], dist[2]));
        t = fmin(t + fmax(1.0f, std::floor(distance / stepSize) + 1.0f) * stepSize, curExit);
        continue;
      }
      shades++;
      Vec4 f = reconstruct(bx, cell, (flags & 1) ? P.filter : 0.0f);
      // fireProRenderField (volume-reconstruction.wgsl:100-116): smoke from its own lattice.
      if (coarse) f.x = renderSmokeAt(bx, cell[0], cell[1], cell[2]);
      if (f.x + glowOf(f.z) > 0.002f) {
        // fireProShade (volume-march.wgsl:103-138)
        float glow[3];
        emission(f, glow);
        float density = fmax(0.0f, f.x) * P.optical[0];
        float scat[3] = {0, 0, 0};
        if (density > 0.0f) {
          Vec4 inc = lightSample(bx, cell[0] / lightRatio, cell[1] / lightRatio, cell[2] / lightRatio);
          float irgb[3] = {inc.x, inc.y, inc.z};
          for (int a = 0; a < 3; a++)
            scat[a] = P.smoke[a] * (P.ambient[a] + direct[a] * inc.w + irgb[a] * 0.65f) * density;
        }
        float ext = extinctionOf(f, P);
        float tau = ext * ds;
        float integral = ds * (1.0f - 0.5f * tau + tau * tau / 6.0f);
        if (tau >= 0.001f) integral = (1.0f - fexp(-tau)) / ext;
        for (int a = 0; a < 3; a++) radiance[a] += transmittance * (scat[a] + glow[a]) * integral;
        transmittance *= 1.0f - fmin(1.0f, ext * integral);
      }
      t += ds;
    }
    out[0] = radiance[0];
    out[1] = radiance[1];
    out[2] = radiance[2];
    out[3] = 1.0f - transmittance;
  }

  // The image; counts[0, 1] the slices' sample positions visited and samples shaded,
  // [2, 3] the march's (integer sums: the same at any thread count).
  void image(float *rgba, uint64_t counts[4] = nullptr) const {
    std::vector<float> slices;
    int sw = 0, sh = 0;
    uint64_t v0 = 0, s0 = 0, v1 = 0, s1 = 0;
    if (P.halfRes) {
      sw = std::max(1, (P.width + 1) / 2);
      sh = std::max(1, (P.height + 1) / 2);
      slices.assign((size_t)sw * sh * SLICES * 4, 0.0f);
      // Rows are independent (each texel and pixel is computed alone): the host's threads
      // share them (OpenMP), with the same result at any thread count.
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : v0, s0)
      for (int y = 0; y < sh; y++)
        for (int x = 0; x < sw; x++) slice(x, y, sw, sh, &slices[((size_t)y * sw + x) * 4], v0, s0);
    }
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : v1, s1)
    for (int y = 0; y < P.height; y++)
      for (int x = 0; x < P.width; x++)
        pixel(x, y, &rgba[((size_t)y * P.width + x) * 4], P.halfRes ? slices.data() : nullptr, sw, sh, v1, s1);
    if (counts) {
      counts[0] = v0;
      counts[1] = s0;
      counts[2] = v1;
      counts[3] = s1;
    }
  }
};

// ---- scene lights (scene-lights.wgsl:36-106; SceneLights.ts:127-218) ----------------------
// Every active brick samples its fields on a 4^3 lattice, weighs the samples by the
// luminance of their emission and gives the totals to the anchor nearest the brick's
// center (height above an anchor counts a quarter; ties to the lower index). The totals
// per anchor are summed in a fixed order: a 64-lane pairwise tree per brick, a 256-lane tree
// per block of 256 consecutive active bricks (others' bricks adding zero), then per lane
// the blocks j, j + 256, ... in order and a 256-lane tree. The reference adds with float
// atomics in scheduling order; this order is the library's.
inline void treeSum(std::vector<float> &v) {
  for (size_t reach = v.size() / 2; reach > 0; reach >>= 1)
    for (size_t lane = 0; lane < reach; lane++) v[lane] += v[lane + reach];
}
inline void sceneLights(const Sim<float> &sim, const Renderer &r, const std::vector<std::array<float, 3>> &anchors,
                        float *sums) {
  const int count = (int)anchors.size();
  std::fill(sums, sums + (size_t)8 * count, 0.0f);
  const int n = (int)sim.activeList.size();
  if (count == 0 || n == 0 || sim.renderBoxes.empty()) return;
  const RenderParams &P = r.P;
  const int B = sim.B, stride = B >> 2;
  std::vector<float> partials((size_t)n * 8, 0.0f);
  std::vector<int> brickAnchor(n, -1);
  for (int bi = 0; bi < n; bi++) {
    const int id = sim.activeList[bi];
    const int page = sim.pages[id];
    if (page < 0) continue;
    const auto &tc = sim.dir.snapshot[id >> 6];
    const int l = id & 63;
    const int brick[3] = {tc[0] * 4 + (l & 3), tc[1] * 4 + ((l >> 2) & 3), tc[2] * 4 + (l >> 4)};
    const float center[3] = {((float)brick[0] + 0.5f) * P.brickSize, ((float)brick[1] + 0.5f) * P.brickSize,
                             ((float)brick[2] + 0.5f) * P.brickSize};
    float nearD[64];
    int nearI[64];
    float lum[64], rgbs[64][3];
    int locals[64][3];
    for (int lane = 0; lane < 64; lane++) {
      const int local[3] = {(lane & 3) * stride + stride / 2, ((lane >> 2) & 3) * stride + stride / 2,
                            (lane >> 4) * stride + stride / 2};
      for (int a = 0; a < 3; a++) locals[lane][a] = local[a];
      const float *fp = &sim.F[((size_t)sim.pageSlot(page) * sim.nF + ((local[2] * B + local[1]) * B + local[0])) * 4];
      Vec4 f{fp[0], fp[1], fp[2], fp[3]};
      if (sim.coarse()) {
        // The coarse density sampled at the lattice point (scene-lights.wgsl:54-61).
        const float ratio = (float)(B / sim.bs);
        float q[3];
        for (int a = 0; a < 3; a++)
          q[a] = (float)(brick[a] * sim.bs) + ((float)local[a] + 0.5f) / ratio;
        f.x = sim.sampleSmoke(sim.S, q);
      }
      r.emission(f, rgbs[lane]);
      lum[lane] = rgbs[lane][0] * 0.2126f + rgbs[lane][1] * 0.7152f + rgbs[lane][2] * 0.0722f;
      float bestD = 1e38f;
      int bestI = 0;
      for (int i = lane; i < count; i += 64) {
        const float dx = center[0] - anchors[i][0], dy = center[1] - anchors[i][1], dz = center[2] - anchors[i][2];
        const float rise = dy > 0.0f ? dy * 0.25f : dy;
        const float d = dx * dx + dz * dz + rise * rise;
        if (d < bestD) {
          bestD = d;
          bestI = i;
        }
      }
      nearD[lane] = bestD;
      nearI[lane] = bestI;
    }
    for (int reach = 32; reach > 0; reach >>= 1)
      for (int lane = 0; lane < reach; lane++) {
        const float od = nearD[lane + reach];
        const int oi = nearI[lane + reach];
        if (od < nearD[lane] || (od == nearD[lane] && oi < nearI[lane])) {
          nearD[lane] = od;
          nearI[lane] = oi;
        }
      }
    const int anchor = nearI[0];
    std::vector<float> tot[7];
    for (auto &t : tot) t.resize(64);
    for (int lane = 0; lane < 64; lane++) {
      float p[3];
      for (int a = 0; a < 3; a++) p[a] = ((float)(brick[a] * B + locals[lane][a]) + 0.5f) * P.voxel;
      tot[0][lane] = lum[lane];
      for (int a = 0; a < 3; a++) tot[1 + a][lane] = (p[a] - anchors[anchor][a]) * lum[lane];
      for (int a = 0; a < 3; a++) tot[4 + a][lane] = rgbs[lane][a];
    }
    for (int k = 0; k < 7; k++) {
      treeSum(tot[k]);
      partials[(size_t)bi * 8 + k] = tot[k][0];
    }
    brickAnchor[bi] = anchor;
  }
  const int BLOCK = 256;
  const int blocks = (n + BLOCK - 1) / BLOCK;
  std::vector<float> blockSums((size_t)blocks * count * 8, 0.0f), red(BLOCK);
  for (int b = 0; b < blocks; b++)
    for (int a = 0; a < count; a++)
      for (int k = 0; k < 7; k++) {
        for (int lane = 0; lane < BLOCK; lane++) {
          const int bi = b * BLOCK + lane;
          red[lane] = bi < n && brickAnchor[bi] == a ? partials[(size_t)bi * 8 + k] : 0.0f;
        }
        treeSum(red);
        blockSums[((size_t)b * count + a) * 8 + k] = red[0];
      }
  for (int a = 0; a < count; a++)
    for (int k = 0; k < 7; k++) {
      for (int lane = 0; lane < BLOCK; lane++) {
        float acc = 0.0f;
        for (int b = lane; b < blocks; b += BLOCK) acc += blockSums[((size_t)b * count + a) * 8 + k];
        red[lane] = acc;
      }
      treeSum(red);
      sums[(size_t)a * 8 + k] = red[0];
    }
}

// VolumeLighting.encode's schedule (VolumeLighting.ts:249-375), as state carried between
// renders of one simulation.
struct RenderCache {
  int version = -1;
  bool stale = true;
  int64_t lastStep = -1, lastLightStep = -1, occupancyStep = -1;
  std::string lastKey;
  int poolSlots = -1, lightDivisor = -1;
  std::vector<RBox> boxes;
  std::vector<int32_t> table;
  std::vector<float> lut = blackbodyTable();
  std::vector<float> R, L, RS;
  std::vector<uint8_t> O;

  // SceneVolumeRenderer.invalidate (SceneVolumeRenderer.ts:254-257) on a reset.
  void invalidate() { lastKey.clear(); }

  // settingsKey (VolumeLighting.ts:262-273) with the scene key of the light direction and
  // whether it shines (SceneVolumeRenderer.ts:417-422).
  static std::string key(const RenderParams &p) {
    char buf[512];
    snprintf(buf, sizeof buf, "%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%d", p.optical[0], p.optical[2],
             p.optical[1], p.flameOpacity, p.fire[3], p.fire[0], p.fire[1], p.fire[2], p.smoke[0], p.smoke[1],
             p.smoke[2], p.sootGlow, p.light[0], p.light[1], p.light[2], p.light[3] > 0.0f ? 1 : 0);
    return buf;
  }

  Renderer view(const Sim<float> &sim, const RenderParams &p) {
    return Renderer(sim, p, boxes, table, lut, R, L, O, RS);
  }

  // Brings the page tables and pools up to the simulation's latest step; returns the
  // light work that ran (Renderer::LightWork), or -1 when encode had nothing to do.
  int prepare(const Sim<float> &sim, const RenderParams &p) {
    Renderer r = view(sim, p);
    if (sim.renderVersion != version) {
      version = sim.renderVersion;
      r.syncBoxes();
      stale = true;
    }
    int B = sim.B, E = B + 1, LE = (1 << p.lightShift) + 1, nO = (B / 8) * (B / 8) * (B / 8);
    bool resized = sim.poolSlots != poolSlots || p.lightDivisor != lightDivisor;
    if (resized) {
      poolSlots = sim.poolSlots;
      lightDivisor = p.lightDivisor;
      R.assign((size_t)poolSlots * E * E * E * 4, 0.0f);
      L.assign((size_t)poolSlots * LE * LE * LE * 4, 0.0f);
      O.assign((size_t)poolSlots * nO, 0);
      if (sim.coarse()) RS.assign((size_t)poolSlots * (sim.bs + 1) * (sim.bs + 1) * (sim.bs + 1), 0.0f);
    }
    bool changed = stale;
    std::string k = key(p);
    bool stepped = sim.steps != lastStep;
    bool rebuild = changed || resized;
    bool smooth = stepped || rebuild;
    bool pages = stepped || changed;
    bool occupancy = occupancyStep != sim.steps || rebuild;
    bool lightDue = sim.steps < lastLightStep || sim.steps - lastLightStep >= 2;
    bool buildLight = (stepped && lightDue) || rebuild || k != lastKey;
    bool fresh = stepped && !buildLight;
    if (boxes.empty() || !(smooth || pages || occupancy || buildLight || fresh)) return -1;
    int work = buildLight ? Renderer::LIGHT_ALL : fresh ? Renderer::LIGHT_FRESH : Renderer::LIGHT_NONE;
    r.prepare(pages, smooth, occupancy, work);
    if (occupancy) occupancyStep = sim.steps;
    if (buildLight) {
      lastLightStep = sim.steps;
      lastKey = k;
    }
    lastStep = sim.steps;
    stale = false;
    return work;
  }
};

}  // namespace oracle
