// Render preparation and the camera ray-march over the solver's bricks.
//
// The renderer marches boxes of bricks (one per cluster of tiles with something to draw);
// each box has a page table saying which of its bricks have pool slots. Preparation, as
// VolumeLighting.encode schedules it (VolumeLighting.ts:249-375), fills per slot:
//  * a render field: smoke smoothed by [1,2,1]/4 per axis, heat/lifetime/fuel as they are,
//    at (B+1)^3 texels, the far apron computed like the interior (smoothSmoke);
//  * occupancy: per 8^3-cell block, whether anything visible or any flame lies within two
//    cells (buildOccupancy);
//  * incident light at the lighting lattice, (B/ld+1)^3 texels with apron: the sun's
//    transmittance to the box's edge and a six-direction glow estimate (buildLighting).
// The march then integrates emission, absorption and single scattering per pixel
// (volume-march.wgsl). Every filtered fetch is an exact FP32 trilinear sum; expression
// trees follow the reference's WGSL term by term with no contraction.
#include "kernels.hpp"

namespace fp {

constexpr int SHADOW_STEPS = 32;  // types.ts SHADOW_STEPS
constexpr int LUT_SIZE = 512;     // blackbody.ts BLACKBODY_TABLE_SIZE
constexpr int SLICES = 8;         // VolumeSlices.ts DEPTH_SLICES

struct RBox {
  int ox, oy, oz, first, sx, sy, sz;
};

__device__ __forceinline__ RBox loadBox(const RenderArgs &r, int i) {
  int4 a = r.boxes[2 * i], b = r.boxes[2 * i + 1];
  return RBox{a.x, a.y, a.z, a.w, b.x, b.y, b.z};
}

// fireProPage (volume-bricks.wgsl:25-34): the page of brick (x, y, z), -1 outside the box
// or without a slot.
__device__ __forceinline__ int pageAt(const RenderArgs &r, const RBox &bx, int x, int y, int z) {
  int lx = x - bx.ox, ly = y - bx.oy, lz = z - bx.oz;
  if (lx < 0 || ly < 0 || lz < 0 || lx >= bx.sx || ly >= bx.sy || lz >= bx.sz) return -1;
  return r.table[bx.first + (lz * bx.sy + ly) * bx.sx + lx];
}

__device__ __forceinline__ float glowOf(float life) { return fminv(fmaxv(life * 5.0f, 0.0f), 1.0f); }
__device__ __forceinline__ float extinctionOf(const float4 &f, const RenderParams &P) {
  return fmaxv(0.0f, f.x) * P.optical[0] + glowOf(f.z) * P.flameOpacity;
}

// fireProLoad over the simulation's fields: the cell through the box's page table.
__device__ __forceinline__ float4 fieldCell(const RenderArgs &r, const RBox &bx, int x, int y, int z) {
  int pg = pageAt(r, bx, x >> r.sF, y >> r.sF, z >> r.sF);
  if (pg < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
  int m = r.B - 1;
  return r.F[(size_t)pageToSlot(pg, r.side) * r.nF + (((z & m) * r.B + (y & m)) * r.B + (x & m))];
}

__device__ __forceinline__ const float4 *texel(const float4 *pool, int slot, int edge, int x, int y, int z) {
  return pool + ((size_t)slot * edge * edge * edge + ((size_t)z * edge + y) * edge + x);
}

// fireProField over the render field (volume-bricks.wgsl:91-104 and :56-88).
__device__ __forceinline__ float4 renderSample(const RenderArgs &r, const RBox &bx, float px, float py, float pz) {
  const float sz = (float)(1 << r.sF);
  float q[3] = {px, py, pz};
  const int o[3] = {bx.ox, bx.oy, bx.oz}, s[3] = {bx.sx, bx.sy, bx.sz};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float lo = (float)o[a] * sz + 0.5f;
    float hi = (float)(o[a] + s[a]) * sz - 0.5f;
    q[a] = fminv(fmaxv(q[a], lo), hi);
  }
  int base[3], brick[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float g = floorf(q[a] - 0.5f);
    base[a] = (int)g;
    brick[a] = base[a] >> r.sF;
    f[a] = (q[a] - 0.5f) - g;
  }
  int pg = pageAt(r, bx, brick[0], brick[1], brick[2]);
  const int last = r.B - 1;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (pg < 0 && (base[0] & last) < last && (base[1] & last) < last && (base[2] & last) < last) return v;
  const int slot = pg >= 0 ? pageToSlot(pg, r.side) : -1;
#pragma unroll
  for (int z = 0; z < 2; z++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int x = 0; x < 2; x++) {
        float wx = x ? f[0] : 1.0f - f[0], wy = y ? f[1] : 1.0f - f[1], wz = z ? f[2] : 1.0f - f[2];
        float w = wx * wy * wz;
        float4 t;
        if (slot >= 0) {
          t = *texel(r.R, slot, r.E, base[0] + x - brick[0] * r.B, base[1] + y - brick[1] * r.B,
                     base[2] + z - brick[2] * r.B);
        } else {
          int cx = base[0] + x, cy = base[1] + y, cz = base[2] + z;
          int p2 = pageAt(r, bx, cx >> r.sF, cy >> r.sF, cz >> r.sF);
          if (p2 < 0) continue;
          t = *texel(r.R, pageToSlot(p2, r.side), r.E, cx & last, cy & last, cz & last);
        }
        v.x += t.x * w;
        v.y += t.y * w;
        v.z += t.z * w;
        v.w += t.w * w;
      }
  return v;
}

// ---- coarse smoke (smokeDivisor 2 or 4) ----
// The simulation's coarse density at smoke cell (x, y, z) through the box's page table
// (fireProLoad on the smoke lattice), zero without a slot.
__device__ __forceinline__ float rawSmokeCell(const RenderArgs &r, const RBox &bx, int x, int y, int z) {
  int pg = pageAt(r, bx, x >> r.sS, y >> r.sS, z >> r.sS);
  if (pg < 0) return 0.f;
  const int m = r.bs - 1;
  return r.S[(size_t)pageToSlot(pg, r.side) * r.nS + (((z & m) * r.bs + (y & m)) * r.bs + (x & m))];
}
// A texel of the smoothed render smoke, slot-major with one-texel aprons.
__device__ __forceinline__ float &smokeTexel(const RenderArgs &r, int slot, int x, int y, int z) {
  return r.RS[(size_t)slot * r.nRS + ((size_t)z * r.ES + y) * r.ES + x];
}
// fireProField on the smoke lattice (volume-bricks.wgsl:56-104), position in smoke cells:
// over the simulation's density (`raw`, whose slots' aprons hold their neighbors' cells)
// or over the smoothed render smoke, whose slots carry computed aprons.
__device__ __forceinline__ float smokeField(const RenderArgs &r, const RBox &bx, float px, float py, float pz,
                                            bool raw) {
  const float sz = (float)r.bs;
  float q[3] = {px, py, pz};
  const int o[3] = {bx.ox, bx.oy, bx.oz}, s[3] = {bx.sx, bx.sy, bx.sz};
#pragma unroll
  for (int a = 0; a < 3; a++) q[a] = fminv(fmaxv(q[a], (float)o[a] * sz + 0.5f), (float)(o[a] + s[a]) * sz - 0.5f);
  int base[3], brick[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float g = floorf(q[a] - 0.5f);
    base[a] = (int)g;
    brick[a] = base[a] >> r.sS;
    f[a] = (q[a] - 0.5f) - g;
  }
  const int pg = pageAt(r, bx, brick[0], brick[1], brick[2]);
  const int last = r.bs - 1;
  if (pg < 0 && (base[0] & last) < last && (base[1] & last) < last && (base[2] & last) < last) return 0.f;
  const int slot = pg >= 0 ? pageToSlot(pg, r.side) : -1;
  float v = 0.f;
#pragma unroll
  for (int z = 0; z < 2; z++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int x = 0; x < 2; x++) {
        float wx = x ? f[0] : 1.0f - f[0], wy = y ? f[1] : 1.0f - f[1], wz = z ? f[2] : 1.0f - f[2];
        float w = wx * wy * wz;
        const int c[3] = {base[0] + x, base[1] + y, base[2] + z};
        float t;
        if (raw) {
          t = rawSmokeCell(r, bx, c[0], c[1], c[2]);
        } else if (slot >= 0) {
          t = smokeTexel(r, slot, c[0] - brick[0] * r.bs, c[1] - brick[1] * r.bs, c[2] - brick[2] * r.bs);
        } else {
          int p2 = pageAt(r, bx, c[0] >> r.sS, c[1] >> r.sS, c[2] >> r.sS);
          if (p2 < 0) continue;
          t = smokeTexel(r, pageToSlot(p2, r.side), c[0] & last, c[1] & last, c[2] & last);
        }
        v += t * w;
      }
  return v;
}
// The smoothed smoke at a field position (renderSmokeAt, volume-common.wgsl:74-78).
__device__ __forceinline__ float renderSmokeAt(const RenderArgs &r, const RBox &bx, float fx, float fy, float fz) {
  const float ratio = (float)(r.B / r.bs);
  return smokeField(r, bx, fx / ratio, fy / ratio, fz / ratio, false);
}

// sampleWorld (volume-common.wgsl:41-51).
__device__ __forceinline__ float4 sampleWorld(const RenderArgs &r, const RBox &bx, float px, float py, float pz) {
  const float size = r.P.brickSize;
  if (px < (float)bx.ox * size || py < (float)bx.oy * size || pz < (float)bx.oz * size ||
      px > (float)(bx.ox + bx.sx) * size || py > (float)(bx.oy + bx.sy) * size || pz > (float)(bx.oz + bx.sz) * size)
    return make_float4(0.f, 0.f, 0.f, 0.f);
  const float fx = px / r.P.voxel, fy = py / r.P.voxel, fz = pz / r.P.voxel;
  float4 f = renderSample(r, bx, fx, fy, fz);
  if (r.coarse) f.x = renderSmokeAt(r, bx, fx, fy, fz);
  return f;
}

// fireProEmission (fire-emission.wgsl:22-56), smoke = (albedo, soot glow).
__device__ __forceinline__ void emission(const RenderArgs &r, const float4 &f, float out[3]) {
  const RenderParams &P = r.P;
  out[0] = out[1] = out[2] = 0.0f;
  if (P.optical[1] <= 0.0f) return;
  float dens = fmaxv(0.0f, f.x);
  float flame = glowOf(f.z) * P.flameOpacity;
  float source[3];
#pragma unroll
  for (int c = 0; c < 3; c++) source[c] = ((1.0f - P.smoke[c]) * dens) * P.optical[0] * P.sootGlow + flame;
  if (source[0] == 0.0f && source[1] == 0.0f && source[2] == 0.0f) return;
  float heat = fmaxv(0.0f, f.y);
  float kelvin = 800.0f + heat / (heat + 2.0f) * (P.fire[3] - 800.0f);
  // The blackbody texel pair, linearly filtered (fire-emission.wgsl:43-45).
  float u = fminv(fmaxv((kelvin - 500.0f) / 9500.0f, 0.0f), 1.0f) * (float)(LUT_SIZE - 1);
  float i0f = floorf(u);
  float fr = u - i0f;
  int i0 = (int)i0f, i1 = i0 + 1 < LUT_SIZE - 1 ? i0 + 1 : LUT_SIZE - 1;
  float4 a = r.lut[i0], b = r.lut[i1];
  float s0 = a.x * (1.0f - fr) + b.x * fr, s1 = a.y * (1.0f - fr) + b.y * fr, s2 = a.z * (1.0f - fr) + b.z * fr,
        s3 = a.w * (1.0f - fr) + b.w * fr;
  float scale = fexp2(fminv(fmaxv(s3, -24.0f), 10.0f));
  float thermal[3] = {s0 * scale, s1 * scale, s2 * scale};
  float tintPeak = fmaxv(P.fire[0], fmaxv(P.fire[1], P.fire[2]));
  float saturation = 1.0f - fminv(P.fire[0], fminv(P.fire[1], P.fire[2])) / fmaxv(tintPeak, 0.000001f);
  float peak = fmaxv(thermal[0], fmaxv(thermal[1], thermal[2]));
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float colored = mixv(thermal[c] * tintPeak, P.fire[c] * peak, saturation);
    out[c] = colored * source[c] * P.optical[1];
  }
}

__device__ __forceinline__ float integratedTransmission(float ext, float distance, float att) {
  float depth = ext * distance;
  if (depth < 0.001f) return distance * (1.0f - 0.5f * depth + depth * depth / 6.0f);
  return (1.0f - att) / ext;
}

__device__ __forceinline__ void normalize3(const float v[3], float out[3]) {
  float len = length3(v[0], v[1], v[2]);
  out[0] = v[0] / len;
  out[1] = v[1] / len;
  out[2] = v[2] / len;
}

// exitDistance (lighting.wgsl:258-266).
__device__ __forceinline__ float exitDistance(const RenderArgs &r, const RBox &bx, const float p[3], const float d[3]) {
  const int o[3] = {bx.ox, bx.oy, bx.oz}, s[3] = {bx.sx, bx.sy, bx.sz};
  float dist[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float edge = (float)(d[a] > 0.0f ? o[a] + s[a] : o[a]) * r.P.brickSize;
    bool moving = fabsv(d[a]) > 0.000001f;
    dist[a] = moving ? (edge - p[a]) / d[a] : 1e30f;
  }
  return fmaxv(0.0f, fminv(dist[0], fminv(dist[1], dist[2])));
}

__constant__ float GLOW_DIRS[6][3] = {{.36f, .48f, -.8f}, {-.36f, -.48f, .8f}, {-.8f, .6f, 0.0f},
                                      {.8f, -.6f, 0.0f},  {.48f, .64f, .6f},   {-.48f, -.64f, -.6f}};

// incidentLight (lighting.wgsl:251-300).
__device__ __forceinline__ float4 incidentLight(const RenderArgs &r, const RBox &bx, const float p[3]) {
  const RenderParams &P = r.P;
  float4 lf = sampleWorld(r, bx, p[0], p[1], p[2]);
  if (lf.x + glowOf(lf.z) < 0.00001f) return make_float4(0.f, 0.f, 0.f, 1.f);
  float sun[3];
  normalize3(P.light, sun);
  float ds = exitDistance(r, bx, p, sun) / (float)SHADOW_STEPS;
  float depth = 0.0f;
  if (P.light[3] > 0.0f) {
    for (int i = 0; i < SHADOW_STEPS; i++) {
      float t = ((float)i + 0.5f) * ds;
      float4 f = sampleWorld(r, bx, p[0] + sun[0] * t, p[1] + sun[1] * t, p[2] + sun[2] * t);
      depth += (fmaxv(0.0f, f.x) * P.optical[2] + glowOf(f.z) * P.flameOpacity) * ds;
      if (depth > 9.0f) break;
    }
  }
  float glow[3] = {0.f, 0.f, 0.f};
  if (P.optical[1] > 0.0f && lf.x > 0.00001f) {
    for (int d = 0; d < 6; d++) {
      float t = 0.0f, interval = 0.075f, transmission = 1.0f;
      for (int s = 0; s < 6; s++) {
        float at = t + interval * 0.5f;
        float4 f = sampleWorld(r, bx, p[0] + GLOW_DIRS[d][0] * at, p[1] + GLOW_DIRS[d][1] * at,
                               p[2] + GLOW_DIRS[d][2] * at);
        float ext = extinctionOf(f, P);
        float att = fexp(-ext * interval);
        float e[3];
        emission(r, f, e);
        float it = integratedTransmission(ext, interval, att);
#pragma unroll
        for (int c = 0; c < 3; c++) glow[c] += transmission * e[c] * it / 6.0f;
        transmission *= att;
        t += interval;
        interval *= 1.65f;
      }
    }
  }
  return make_float4(glow[0], glow[1], glow[2], fexp(-depth));
}

// A brick of the active list with a slot inside a box (lighting.wgsl:34-60).
__device__ __forceinline__ bool slotted(const RenderArgs &r, int id, int &page, RBox &bx, int brick[3]) {
  page = r.pages[id];
  int4 tile = r.tiles[id >> 6];
  if (page < 0 || tile.w < 0) return false;
  bx = loadBox(r, tile.w);
  int l = id & 63;
  brick[0] = tile.x * 4 + (l & 3);
  brick[1] = tile.y * 4 + ((l >> 2) & 3);
  brick[2] = tile.z * 4 + (l >> 4);
  int lx = brick[0] - bx.ox, ly = brick[1] - bx.oy, lz = brick[2] - bx.oz;
  return lx >= 0 && ly >= 0 && lz >= 0 && lx < bx.sx && ly < bx.sy && lz < bx.sz;
}

// exportPages (lighting.wgsl:10-32): a lane per brick id.
__global__ void __launch_bounds__(256) k_export_pages(RenderArgs r) {
  const int id = blockIdx.x * 256 + threadIdx.x;
  if (id >= r.brickCount) return;
  int4 tile = r.tiles[id >> 6];
  if (tile.w < 0) return;
  RBox bx = loadBox(r, tile.w);
  int l = id & 63;
  int lx = tile.x * 4 + (l & 3) - bx.ox, ly = tile.y * 4 + ((l >> 2) & 3) - bx.oy, lz = tile.z * 4 + (l >> 4) - bx.oz;
  if (lx < 0 || ly < 0 || lz < 0 || lx >= bx.sx || ly >= bx.sy || lz >= bx.sz) return;
  r.table[bx.first + (lz * bx.sy + ly) * bx.sx + lx] = r.pages[id];
}

// smoothSmoke (lighting.wgsl:62-97). A workgroup per chunk of a listed brick's texels.
// With `staged`, the smoke of cells [brick*B - 1, brick*B + B + 1] per axis, read through
// the page table, sits in LDS (B <= 16): every sample position of the eight half-cell
// taps, clamped to the box, has its two cells per axis there. The pages of the 27 bricks
// around are staged either way.
__global__ void __launch_bounds__(256) k_smooth_smoke(RenderArgs r, int chunks, int staged) {
  const int item = xcdItem((int)blockIdx.x, r.counts[0] * chunks);
  if (item < 0) return;
  const int i = item / chunks, chunk = item % chunks;
  const int id = r.active[i];
  int page, brick[3];
  RBox bx;
  if (!slotted(r, id, page, bx, brick)) return;
  extern __shared__ float cells[];
  __shared__ int near[27];
  const int B = r.B, sF = r.sF, last = B - 1;
  const int W = B + 3;  // staged edge
  if (threadIdx.x < 27) {
    int t = threadIdx.x;
    near[t] = pageAt(r, bx, brick[0] + t % 3 - 1, brick[1] + (t / 3) % 3 - 1, brick[2] + t / 9 - 1);
  }
  __syncthreads();
  const int x0 = brick[0] * B - 1, y0 = brick[1] * B - 1, z0 = brick[2] * B - 1;
  // The smoke of cell (cx, cy, cz) through the page table: zero without a slot.
  auto smoke = [&](int cx, int cy, int cz) -> float {
    int pg = near[((cz >> sF) - brick[2] + 1) * 9 + ((cy >> sF) - brick[1] + 1) * 3 + ((cx >> sF) - brick[0] + 1)];
    return pg < 0 ? 0.0f
                  : r.F[(size_t)pageToSlot(pg, r.side) * r.nF + (((cz & last) * B + (cy & last)) * B + (cx & last))].x;
  };
  if (staged) {
    for (int c = threadIdx.x; c < W * W * W; c += 256) cells[c] = smoke(x0 + c % W, y0 + (c / W) % W, z0 + c / (W * W));
    __syncthreads();
  }
  const int slot = pageToSlot(page, r.side);
  const int E = r.E;
  const float sz = (float)B;
  float lo[3], hi[3];
  const int o[3] = {bx.ox, bx.oy, bx.oz}, s[3] = {bx.sx, bx.sy, bx.sz};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    lo[a] = (float)o[a] * sz + 0.5f;
    hi[a] = (float)(o[a] + s[a]) * sz - 0.5f;
  }
  const int first[3] = {brick[0] * B, brick[1] * B, brick[2] * B};
  for (int t = chunk * 256 + threadIdx.x; t < r.nR; t += chunks * 256) {
    int lx = t % E, ly = (t / E) % E, lz = t / (E * E);
    int cell[3] = {first[0] + lx, first[1] + ly, first[2] + lz};
    float4 original = fieldCell(r, bx, cell[0], cell[1], cell[2]);
    if (r.coarse) {
      // Smoke lives on its own grid: the render field is the fields as they are
      // (VolumeLighting.ts:206-208), with aprons.
      r.R[(size_t)slot * r.nR + t] = original;
      continue;
    }
    float density = 0.0f;
    for (int corner = 0; corner < 8; corner++) {
      float q[3] = {(float)cell[0] + (float)(corner & 1), (float)cell[1] + (float)((corner >> 1) & 1),
                    (float)cell[2] + (float)((corner >> 2) & 1)};
      int base[3];
      float f[3];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        q[a] = fminv(fmaxv(q[a], lo[a]), hi[a]);
        float g = floorf(q[a] - 0.5f);
        base[a] = (int)g;
        f[a] = (q[a] - 0.5f) - g;
      }
      int nb = ((base[2] >> sF) - brick[2] + 1) * 9 + ((base[1] >> sF) - brick[1] + 1) * 3 + ((base[0] >> sF) - brick[0] + 1);
      // A tap whose base brick has no slot and that does not reach the next brick is zero.
      if (near[nb] < 0 && (base[0] & last) < last && (base[1] & last) < last && (base[2] & last) < last) continue;
      float v = 0.0f;
#pragma unroll
      for (int z = 0; z < 2; z++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
          for (int x = 0; x < 2; x++) {
            float wx = x ? f[0] : 1.0f - f[0], wy = y ? f[1] : 1.0f - f[1], wz = z ? f[2] : 1.0f - f[2];
            float c = staged ? cells[((base[2] + z - z0) * W + (base[1] + y - y0)) * W + (base[0] + x - x0)]
                             : smoke(base[0] + x, base[1] + y, base[2] + z);
            v += c * (wx * wy * wz);
          }
      density += v;
    }
    r.R[(size_t)slot * r.nR + t] = make_float4(fmaxv(0.0f, density * 0.125f), original.y, original.z, original.w);
  }
}

// smoothCoarseSmoke (lighting.wgsl:105-129): the [1,2,1]/4 filter per axis on the smoke
// lattice, as eight half-cell taps of the simulation's density, at every texel of a listed
// brick's render smoke slot, the far apron computed like the interior. A workgroup per
// listed brick.
__global__ void __launch_bounds__(256) k_smooth_coarse(RenderArgs r) {
  const int i = xcdItem((int)blockIdx.x, r.counts[0]);
  if (i < 0) return;
  int page, brick[3];
  RBox bx;
  if (!slotted(r, r.active[i], page, bx, brick)) return;
  const int slot = pageToSlot(page, r.side), ES = r.ES;
  for (int t = threadIdx.x; t < r.nRS; t += blockDim.x) {
    const int cell[3] = {brick[0] * r.bs + t % ES, brick[1] * r.bs + (t / ES) % ES, brick[2] * r.bs + t / (ES * ES)};
    float density = 0.0f;
    for (int corner = 0; corner < 8; corner++)
      density += smokeField(r, bx, (float)cell[0] + (float)(corner & 1), (float)cell[1] + (float)((corner >> 1) & 1),
                            (float)cell[2] + (float)((corner >> 2) & 1), true);
    r.RS[(size_t)slot * r.nRS + t] = fmaxv(0.0f, density * 0.125f);
  }
}

// cachedSmokeCell (lighting.wgsl:133-152): the smoothed smoke of a smoke cell; for a
// brick without a slot, the apron texel a slot before it computed for the cell.
__device__ __forceinline__ float cachedSmokeCell(const RenderArgs &r, const RBox &bx, int x, int y, int z) {
  const int n = r.bs, m = n - 1, sh = r.sS;
  const int b[3] = {x >> sh, y >> sh, z >> sh}, l[3] = {x & m, y & m, z & m};
  const int pg = pageAt(r, bx, b[0], b[1], b[2]);
  if (pg >= 0) return smokeTexel(r, pageToSlot(pg, r.side), l[0], l[1], l[2]);
  for (int mk = 1; mk < 8; mk++) {
    const int st[3] = {mk & 1, (mk >> 1) & 1, (mk >> 2) & 1};
    if ((st[0] && l[0]) || (st[1] && l[1]) || (st[2] && l[2])) continue;
    const int prev = pageAt(r, bx, b[0] - st[0], b[1] - st[1], b[2] - st[2]);
    if (prev >= 0)
      return smokeTexel(r, pageToSlot(prev, r.side), l[0] + st[0] * n, l[1] + st[1] * n, l[2] + st[2] * n);
  }
  return 0.0f;
}

// buildOccupancy (lighting.wgsl:150-238): per 8^3 block, the 12^3 cells of the block and
// its two-cell halo whose bricks have slots. A workgroup per listed brick; its lanes OR
// their flags through LDS. With coarse smoke the fine cells give flame only, and a block is
// visible when the largest glow there plus the largest smoothed smoke among the smoke
// donors of the block and halo passes the cutoff.
__global__ void __launch_bounds__(256) k_occupancy(RenderArgs r) {
  const int i = xcdItem((int)blockIdx.x, r.counts[0]);
  if (i < 0) return;
  const int id = r.active[i];
  int page, brick[3];
  RBox bx;
  if (!slotted(r, id, page, bx, brick)) return;
  __shared__ uint32_t flagsOf[64], glowMax[64], smokeMax[64];
  const int B = r.B, sF = r.sF, m = B - 1, ps = r.perSide;
  for (int b = threadIdx.x; b < r.nO; b += 256) {
    flagsOf[b] = 0u;
    glowMax[b] = 0u;
    smokeMax[b] = 0u;
  }
  __syncthreads();
  // 216 units of 2^3 cells per block.
  for (int w = threadIdx.x; w < r.nO * 216; w += 256) {
    int block = w / 216, unit = w % 216;
    int b0 = block % ps, b1 = (block / ps) % ps, b2 = block / (ps * ps);
    int cx = brick[0] * B + b0 * 8 - 2 + (unit % 6) * 2, cy = brick[1] * B + b1 * 8 - 2 + ((unit / 6) % 6) * 2,
        cz = brick[2] * B + b2 * 8 - 2 + (unit / 36) * 2;
    int pg = pageAt(r, bx, cx >> sF, cy >> sF, cz >> sF);
    if (pg < 0) continue;
    int s2 = pageToSlot(pg, r.side);
    uint32_t flags = 0u;
    float glowHigh = 0.0f;
    for (int corner = 0; corner < 8; corner++) {
      float4 f = *texel(r.R, s2, r.E, (cx & m) + (corner & 1), (cy & m) + ((corner >> 1) & 1), (cz & m) + ((corner >> 2) & 1));
      float glow = glowOf(f.z);
      float smoke = r.coarse ? 0.0f : f.x;
      if (glow > 0.002f) flags |= 1u;
      if (smoke + glow > 0.002f) flags |= 2u;
      glowHigh = fmaxv(glowHigh, glow);
    }
    if (flags) atomicOr(&flagsOf[block], flags);
    if (r.coarse && glowHigh > 0.0f) atomicMax(&glowMax[block], __float_as_uint(glowHigh));
  }
  if (r.coarse) {
    // Every trilinear smoke donor of each block and its halo, clamped to the box.
    const float ratio = (float)(B / r.bs);
    const int n = r.bs;
    const int boxLo[3] = {bx.ox * n, bx.oy * n, bx.oz * n};
    const int boxHi[3] = {(bx.ox + bx.sx) * n - 1, (bx.oy + bx.sy) * n - 1, (bx.oz + bx.sz) * n - 1};
    for (int block = 0; block < r.nO; block++) {
      const int b3[3] = {block % ps, (block / ps) % ps, block / (ps * ps)};
      int lo[3], size[3];
      for (int a = 0; a < 3; a++) {
        const int origin = brick[a] * B + b3[a] * 8 - 2;
        lo[a] = (int)floorf((float)origin / ratio - 0.5f);
        const int hi = (int)floorf((float)(origin + 12) / ratio - 0.5f) + 1;
        size[a] = hi - lo[a] + 1;
      }
      float laneSmoke = 0.0f;
      for (int c = threadIdx.x; c < size[0] * size[1] * size[2]; c += 256) {
        int q[3] = {lo[0] + c % size[0], lo[1] + (c / size[0]) % size[1], lo[2] + c / (size[0] * size[1])};
        for (int a = 0; a < 3; a++) q[a] = q[a] < boxLo[a] ? boxLo[a] : q[a] > boxHi[a] ? boxHi[a] : q[a];
        laneSmoke = fmaxv(laneSmoke, cachedSmokeCell(r, bx, q[0], q[1], q[2]));
      }
      if (laneSmoke > 0.0f) atomicMax(&smokeMax[block], __float_as_uint(laneSmoke));
    }
  }
  __syncthreads();
  if (r.coarse)
    for (int b = threadIdx.x; b < r.nO; b += 256)
      if (__uint_as_float(smokeMax[b]) + __uint_as_float(glowMax[b]) > 0.002f) flagsOf[b] |= 2u;
  __syncthreads();
  const int slot = pageToSlot(page, r.side);
  for (int b = threadIdx.x; b < r.nO; b += 256) r.O[(size_t)slot * r.nO + b] = (uint8_t)flagsOf[b];
}

// lightSlot (lighting.wgsl:302-345). A workgroup per listed brick (`fresh` = 0, every
// brick: a texel of the far apron is left to the brick past it when that brick has a
// slot, and texels on the near faces also fill the aprons of the bricks before) or per
// slot given out this step (`fresh` = 1, lightFresh: every texel of the slot).
__global__ void __launch_bounds__(64) k_light(RenderArgs r, int fresh) {
  const int i = xcdItem((int)blockIdx.x, fresh ? r.counts[1] : r.counts[0]);
  if (i < 0) return;
  const int id = fresh ? r.zeroList[2 * i] : r.active[i];
  int page, brick[3];
  RBox bx;
  if (!slotted(r, id, page, bx, brick)) return;
  const int ls = r.lside, LE = r.LE;
  const int slot = pageToSlot(page, r.side);
  for (int t = threadIdx.x; t < r.nL; t += 64) {
    int local[3] = {t % LE, (t / LE) % LE, t / (LE * LE)};
    int past[3] = {local[0] == ls, local[1] == ls, local[2] == ls};
    bool anyPast = past[0] | past[1] | past[2];
    if (!fresh && anyPast && pageAt(r, bx, brick[0] + past[0], brick[1] + past[1], brick[2] + past[2]) >= 0) continue;
    float p[3];
#pragma unroll
    for (int a = 0; a < 3; a++) p[a] = ((float)(brick[a] * ls + local[a]) + 0.5f) * r.P.texel;
    float4 v = incidentLight(r, bx, p);
    r.Lt[(size_t)slot * r.nL + t] = v;
    if (fresh || anyPast) continue;
    for (int mk = 1; mk < 8; mk++) {
      int st[3] = {mk & 1, (mk >> 1) & 1, (mk >> 2) & 1};
      if ((st[0] && local[0]) || (st[1] && local[1]) || (st[2] && local[2])) continue;
      int pg = pageAt(r, bx, brick[0] - st[0], brick[1] - st[1], brick[2] - st[2]);
      if (pg >= 0)
        *const_cast<float4 *>(texel(r.Lt, pageToSlot(pg, r.side), LE, local[0] + st[0] * ls, local[1] + st[1] * ls,
                                    local[2] + st[2] * ls)) = v;
    }
  }
}

// fireProHit (volume-march.wgsl:27-35).
__device__ __forceinline__ void hit(const RenderArgs &r, const float o[3], const float d[3], const RBox &bx, float &entry,
                                    float &exit) {
  const int bo[3] = {bx.ox, bx.oy, bx.oz}, bs[3] = {bx.sx, bx.sy, bx.sz};
  float lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float safe = fabsv(d[a]) > 0.000001f ? d[a] : 0.000001f;
    float u = ((float)bo[a] * r.P.brickSize - o[a]) / safe;
    float v = ((float)(bo[a] + bs[a]) * r.P.brickSize - o[a]) / safe;
    lo[a] = fminv(u, v);
    hi[a] = fmaxv(u, v);
  }
  entry = fmaxv(0.0f, fmaxv(lo[0], fmaxv(lo[1], lo[2])));
  exit = fminv(hi[0], fminv(hi[1], hi[2]));
}

// fireProLight (volume-bricks.wgsl:160-179).
__device__ __forceinline__ float4 lightSample(const RenderArgs &r, const RBox &bx, float px, float py, float pz) {
  const float sz = (float)r.lside;
  float q[3] = {px, py, pz};
  const int o[3] = {bx.ox, bx.oy, bx.oz}, s[3] = {bx.sx, bx.sy, bx.sz};
  int base[3], brick[3];
  float f[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float lo = (float)o[a] * sz + 0.5f;
    q[a] = fminv(fmaxv(q[a], lo), (float)(o[a] + s[a]) * sz - 0.5f);
    float g = floorf(q[a] - 0.5f);
    base[a] = (int)g;
    brick[a] = base[a] >> r.P.lightShift;
    f[a] = (q[a] - 0.5f) - g;
  }
  int pg = pageAt(r, bx, brick[0], brick[1], brick[2]);
  if (pg < 0) return make_float4(0.f, 0.f, 0.f, 1.f);
  const int slot = pageToSlot(pg, r.side);
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int z = 0; z < 2; z++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int x = 0; x < 2; x++) {
        float wx = x ? f[0] : 1.0f - f[0], wy = y ? f[1] : 1.0f - f[1], wz = z ? f[2] : 1.0f - f[2];
        float w = wx * wy * wz;
        float4 t = *texel(r.Lt, slot, r.LE, base[0] + x - brick[0] * r.lside, base[1] + y - brick[1] * r.lside,
                          base[2] + z - brick[2] * r.lside);
        v.x += t.x * w;
        v.y += t.y * w;
        v.z += t.z * w;
        v.w += t.w * w;
      }
  return v;
}

// fireProReconstruct (volume-reconstruction.wgsl:5-96).
__device__ __forceinline__ float4 reconstruct(const RenderArgs &r, const RBox &bx, const float c[3], float mode) {
  if (mode < 0.5f) return renderSample(r, bx, c[0], c[1], c[2]);
  float pos[2][3], wt[2][3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float grid = c[a] - 0.5f;
    float base = floorf(grid);
    float q = grid - base;
    if (mode < 1.5f) {
      float center = floorf(grid + 0.5f);
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
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int k = 0; k < 2; k++) {
        float w = wt[i][0] * wt[j][1] * wt[k][2];
        float4 s = renderSample(r, bx, pos[i][0], pos[j][1], pos[k][2]);
        v.x += w * s.x;
        v.y += w * s.y;
        v.z += w * s.z;
        v.w += w * s.w;
      }
  return make_float4(fmaxv(v.x, 0.0f), fmaxv(v.y, 0.0f), fmaxv(v.z, 0.0f), fmaxv(v.w, 0.0f));
}

// fireProJitter (volume-march.wgsl:92-100).
__device__ __forceinline__ float jitter(uint32_t px, uint32_t py, float time) {
  uint32_t s0 = px * 1664525u + 1013904223u, s1 = py * 1664525u + 1013904223u,
           s2 = (uint32_t)(time * 60.0f) * 1664525u + 1013904223u;
  s0 += s1 * s2;
  s1 += s2 * s0;
  s2 += s0 * s1;
  s0 ^= s0 >> 16u;
  s1 ^= s1 >> 16u;
  s2 ^= s2 >> 16u;
  s0 += s1 * s2;
  return (float)(s0 >> 8u) / 16777216.0f;
}

// fireProSceneVolume's beauty path (scene-volume.wgsl:38-48, :146-167) with the march of
// volume-march.wgsl:183-266: a lane per pixel, 16 x 16 pixel tiles per workgroup.
// fireProRay (volume-march.wgsl:18-24): the ray through screen position (uvx, uvy).
__device__ __forceinline__ void rayOf(const RenderParams &P, float uvx, float uvy, float o[3], float d[3]) {
  float cx = uvx * 2.0f + -1.0f, cy = uvy * -2.0f + 1.0f;
  float nh[4], fh[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    nh[k] = ((P.ivp[k] * cx + P.ivp[4 + k] * cy) + P.ivp[8 + k] * 0.0f) + P.ivp[12 + k] * 1.0f;
    fh[k] = ((P.ivp[k] * cx + P.ivp[4 + k] * cy) + P.ivp[8 + k] * 0.99999f) + P.ivp[12 + k] * 1.0f;
  }
  float dd[3];
#pragma unroll
  for (int a = 0; a < 3; a++) o[a] = nh[a] / nh[3];
#pragma unroll
  for (int a = 0; a < 3; a++) dd[a] = fh[a] / fh[3] - o[a];
  normalize3(dd, d);
}

// A ray's march through the boxes (fireProMarch, volume-march.wgsl:183-266), from t with
// the state (transmittance, radiance) so far. With `count` > 0 it is the half-resolution
// march of buildSlices (volume-slices.wgsl:27-122): steps also stop at each of `count`
// evenly spaced depths from `sx` to `sy`, where the state is stored into `slots`.
struct MarchState {
  float t, transmittance, rad[3];
  int steps;   // loop iterations taken (a cost measure for scheduling)
  int visits;  // sample positions looked up in the occupancy
  int shades;  // of those, samples reconstructed and shaded
};
__device__ __forceinline__ void march(const RenderArgs &r, const float o[3], const float d[3], float stepSize, float jit,
                                      float end, MarchState &st, int count = 0, float sx = 0.f, float sy = 0.f,
                                      float4 *slots = nullptr, size_t stride = 0) {
  const RenderParams &P = r.P;
  // fireProDirect (volume-march.wgsl:138-143)
  float ln[3];
  normalize3(P.light, ln);
  const float g = P.optical[3];
  float mu = d[0] * ln[0] + d[1] * ln[1] + d[2] * ln[2];
  float base = fmaxv(0.05f, 1.0f + g * g - 2.0f * g * mu);
  float phase = (1.0f - g * g) / (base * sqrtf(base));
  float direct[3];
#pragma unroll
  for (int a = 0; a < 3; a++) direct[a] = P.sceneLight[a] * P.light[3] * (0.35f + phase * 0.3f);
  float safe[3], exitSide[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    safe[a] = fabsv(d[a]) > 0.000001f ? d[a] : 0.000001f;
    exitSide[a] = d[a] > 0.0f ? 1.0f : 0.0f;
  }
  float t = st.t, transmittance = st.transmittance, rad[3] = {st.rad[0], st.rad[1], st.rad[2]};
  RBox bx{0, 0, 0, 0, 0, 0, 0};
  float curExit = t;
  const float lightRatio = (float)P.lightDivisor;
  const int m = r.B - 1, ps = r.perSide;
  const bool record = count > 0;
  const float spacing = record ? fmaxv(sy - sx, 0.0f) / (float)count : 0.0f;
  float depth = sx + spacing;
  int recorded = 0;
  int steps = 0, visits = 0, shades = 0;
  while (true) {
    if (record ? (recorded >= count || t >= sy || transmittance < 0.008f) : (t >= end || transmittance < 0.008f)) break;
    steps++;
    if (t >= curExit) {
      // fireProEnter (volume-march.wgsl:160-172)
      float nx = 1e30f, ny = 0.0f;
      int ni = -1;
      for (int i = 0; i < r.boxCount; i++) {
        float en, ex;
        hit(r, o, d, loadBox(r, i), en, ex);
        if (ex > en && ex > t && en < nx) {
          nx = en;
          ny = ex;
          ni = i;
        }
      }
      if (ni < 0) break;
      bx = loadBox(r, ni);
      curExit = ny;
      t = fmaxv(t, nx);
      if (!record) continue;
    } else {
      visits++;
      float ds = fminv(stepSize, fminv(record ? depth : end, curExit) - t);
      float at = t + jit * ds;
      float p[3] = {o[0] + d[0] * at, o[1] + d[1] * at, o[2] + d[2] * at};
      float cell[3] = {p[0] / P.voxel, p[1] / P.voxel, p[2] / P.voxel};
      int c[3] = {(int)floorf(cell[0]), (int)floorf(cell[1]), (int)floorf(cell[2])};
      int pg = pageAt(r, bx, c[0] >> P.shift, c[1] >> P.shift, c[2] >> P.shift);
      float size = P.brickSize;
      uint32_t flags = 0u;
      if (pg >= 0) {
        int blk = (((c[2] & m) >> 3) * ps + ((c[1] & m) >> 3)) * ps + ((c[0] & m) >> 3);
        flags = r.O[(size_t)pageToSlot(pg, r.side) * r.nO + blk];
        size = P.voxel * 8.0f;
      }
      if (!(flags & 2u)) {
        float dist[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
          float edge = (floorf(p[a] / size) + exitSide[a]) * size;
          dist[a] = (edge - p[a]) / safe[a];
        }
        float distance = fminv(dist[0], fminv(dist[1], dist[2]));
        t = fminv(t + fmaxv(1.0f, floorf(distance / stepSize) + 1.0f) * stepSize, curExit);
      } else {
        shades++;
        float4 f = reconstruct(r, bx, cell, (flags & 1u) ? P.filter : 0.0f);
        // Smoke was smoothed once on its own lattice (fireProRenderField, volume-reconstruction.wgsl:100-116).
        if (r.coarse) f.x = renderSmokeAt(r, bx, cell[0], cell[1], cell[2]);
        if (f.x + glowOf(f.z) > 0.002f) {
          // fireProShade (volume-march.wgsl:103-138)
          float glow[3];
          emission(r, f, glow);
          float density = fmaxv(0.0f, f.x) * P.optical[0];
          float scat[3] = {0.f, 0.f, 0.f};
          if (density > 0.0f) {
            float4 inc = lightSample(r, bx, cell[0] / lightRatio, cell[1] / lightRatio, cell[2] / lightRatio);
            float irgb[3] = {inc.x, inc.y, inc.z};
#pragma unroll
            for (int a = 0; a < 3; a++)
              scat[a] = P.smoke[a] * (P.ambient[a] + direct[a] * inc.w + irgb[a] * 0.65f) * density;
          }
          float ext = extinctionOf(f, P);
          float tau = ext * ds;
          float integral = ds * (1.0f - 0.5f * tau + tau * tau / 6.0f);
          if (tau >= 0.001f) integral = (1.0f - fexp(-tau)) / ext;
#pragma unroll
          for (int a = 0; a < 3; a++) rad[a] += transmittance * (scat[a] + glow[a]) * integral;
          transmittance *= 1.0f - fminv(1.0f, ext * integral);
        }
        t += ds;
      }
    }
    if (!record) continue;
    // A skip or a jump between boxes can pass several depths; each keeps the current values.
    while (recorded < count && t >= depth) {
      slots[recorded * stride] = make_float4(rad[0], rad[1], rad[2], 1.0f - transmittance);
      recorded++;
      depth = sx + spacing * (float)(recorded + 1);
    }
  }
  // Depths past the exit or an opaque ray hold the final values.
  for (int k = recorded; k < count; k++) slots[k * stride] = make_float4(rad[0], rad[1], rad[2], 1.0f - transmittance);
  st.t = t;
  st.steps += steps;
  st.visits += visits;
  st.shades += shades;
  st.transmittance = transmittance;
  st.rad[0] = rad[0];
  st.rad[1] = rad[1];
  st.rad[2] = rad[2];
}

// fireProSpan (volume-march.wgsl:39-48): the first entry, the last exit and the length
// inside the boxes.
__device__ __forceinline__ void spanOf(const RenderArgs &r, const float o[3], const float d[3], float &sx, float &sy,
                                       float &sz) {
  sx = 1e30f;
  sy = 0.0f;
  sz = 0.0f;
  for (int i = 0; i < r.boxCount; i++) {
    float en, ex;
    hit(r, o, d, loadBox(r, i), en, ex);
    if (ex > en) {
      sx = fminv(sx, en);
      sy = fmaxv(sy, ex);
      sz = sz + ex - en;
    }
  }
}

// buildSlices (volume-slices.wgsl:27-122): the half-resolution march, a lane per texel of
// a (ceil(w/2), ceil(h/2)) grid, storing its state at SLICES evenly spaced depths from the
// first box entry to the last exit: slices[(k * sh + y) * sw + x].
__device__ __forceinline__ MarchState sliceTexel(const RenderArgs &r, float4 *slices, int sw, int sh, int px, int py) {
  const RenderParams &P = r.P;
  float o[3], d[3], sx, sy, sz;
  rayOf(P, ((float)px + 0.5f) / (float)sw, ((float)py + 0.5f) / (float)sh, o, d);
  spanOf(r, o, d, sx, sy, sz);
  const float stepSize = fmaxv(P.voxel * 0.5f, sz / P.raySteps);
  MarchState st{sx, 1.0f, {0.f, 0.f, 0.f}, 0, 0, 0};
  march(r, o, d, stepSize, jitter((uint32_t)px, (uint32_t)py, P.time), sy, st, SLICES, sx, sy,
        slices + ((size_t)py * sw + px), (size_t)sw * sh);
  return st;
}
// Adds a workgroup's sample counts to stats[0] (positions visited) and stats[1] (samples
// shaded): LDS sums, then one atomic each.
__device__ __forceinline__ void countSamples(unsigned long long *stats, int visits, int shades) {
  __shared__ unsigned int sums[2];
  if (threadIdx.x < 2) sums[threadIdx.x] = 0u;
  __syncthreads();
  if (visits) atomicAdd(&sums[0], (unsigned int)visits);
  if (shades) atomicAdd(&sums[1], (unsigned int)shades);
  __syncthreads();
  if (threadIdx.x < 2 && sums[threadIdx.x]) atomicAdd(&stats[threadIdx.x], (unsigned long long)sums[threadIdx.x]);
}
// The same over 16 x 16 texel tiles, a workgroup each, taken in the order `order` (tile ids
// row-major, longest first from the last frame) or in id order; each tile's cost, the sum
// over its wavefronts of their longest march, goes to `cost`.
__global__ void __launch_bounds__(256) k_slices(RenderArgs r, float4 *slices, int sw, int sh, const int *order,
                                                uint32_t *cost, unsigned long long *stats) {
  __shared__ uint32_t waveCost[4];
  const int tilesX = (sw + 15) >> 4;
  const int tile = order ? order[blockIdx.x] : (int)blockIdx.x;
  const int px = (tile % tilesX) * 16 + (threadIdx.x & 15), py = (tile / tilesX) * 16 + (threadIdx.x >> 4);
  if (threadIdx.x < 4) waveCost[threadIdx.x] = 0u;
  __syncthreads();
  MarchState st{0.f, 1.0f, {0.f, 0.f, 0.f}, 0, 0, 0};
  if (px < sw && py < sh) st = sliceTexel(r, slices, sw, sh, px, py);
  atomicMax(&waveCost[threadIdx.x >> 6], (uint32_t)st.steps);
  __syncthreads();
  if (threadIdx.x == 0 && cost) cost[tile] = waveCost[0] + waveCost[1] + waveCost[2] + waveCost[3];
  if (stats) countSamples(stats, st.visits, st.shades);
}
// Tile ids by decreasing cost (by its binary order of magnitude; ties in any order), one
// workgroup: the next frame starts its longest tiles first, so the short ones fill the tail.
__global__ void __launch_bounds__(1024) k_order_tiles(const uint32_t *cost, int tiles, int *order) {
  __shared__ int start[33];
  if (threadIdx.x < 33) start[threadIdx.x] = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < tiles; t += blockDim.x) atomicAdd(&start[__clz((int)cost[t])], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int sum = 0;
    for (int b = 0; b < 33; b++) {
      const int c = start[b];
      start[b] = sum;
      sum += c;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < tiles; t += blockDim.x) order[atomicAdd(&start[__clz((int)cost[t])], 1)] = t;
}

// fireProSceneVolume's beauty path (scene-volume.wgsl:38-48, :146-167) with the march of
// volume-march.wgsl:183-266: a lane per pixel, 16 x 16 pixel tiles per workgroup. With
// `slices` (half resolution), the march resumes from the last stored depth in front of the
// scene, sampled bilinearly; with no scene depth that is the last one.
__device__ __forceinline__ void marchPixel(const RenderArgs &r, float4 *out, const float4 *slices, int sw, int sh,
                                           int px, int py, int &visits, int &shades) {
  const RenderParams &P = r.P;
  float4 result = make_float4(0.f, 0.f, 0.f, 0.f);
  const float uvx = ((float)px + 0.5f) / (float)P.width, uvy = ((float)py + 0.5f) / (float)P.height;
  float o[3], d[3], sx, sy, sz;
  rayOf(P, uvx, uvy, o, d);
  spanOf(r, o, d, sx, sy, sz);
  if (sx < sy) {
    const float end = sy;
    const float stepSize = fmaxv(P.voxel * 0.5f, sz / P.raySteps);
    MarchState st{sx, 1.0f, {0.f, 0.f, 0.f}, 0, 0, 0};
    if (slices) {
      const float count = (float)SLICES;
      const float slice = floorf((end - sx) / (sy - sx) * count);
      if (slice >= 1.0f) {
        const float k = fminv(slice, count);
        // A linear fetch at (uv, (k - 1/2) / count): one depth layer, bilinear in x and y,
        // clamped to the edge.
        const float fx0 = uvx * (float)sw - 0.5f, fy0 = uvy * (float)sh - 0.5f;
        const float gx = floorf(fx0), gy = floorf(fy0);
        const float fx = fx0 - gx, fy = fy0 - gy;
        const int x0 = min(max((int)gx, 0), sw - 1), x1 = min(max((int)gx + 1, 0), sw - 1);
        const int y0 = min(max((int)gy, 0), sh - 1), y1 = min(max((int)gy + 1, 0), sh - 1);
        const float4 *layer = slices + (size_t)((int)k - 1) * sw * sh;
        const float4 v00 = layer[(size_t)y0 * sw + x0], v10 = layer[(size_t)y0 * sw + x1];
        const float4 v01 = layer[(size_t)y1 * sw + x0], v11 = layer[(size_t)y1 * sw + x1];
        const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
        const float4 sv = make_float4(((v00.x * w00 + v10.x * w10) + v01.x * w01) + v11.x * w11,
                                      ((v00.y * w00 + v10.y * w10) + v01.y * w01) + v11.y * w11,
                                      ((v00.z * w00 + v10.z * w10) + v01.z * w01) + v11.z * w11,
                                      ((v00.w * w00 + v10.w * w10) + v01.w * w01) + v11.w * w11);
        st.t = sx + (sy - sx) * k / count;
        st.transmittance = 1.0f - sv.w;
        st.rad[0] = sv.x;
        st.rad[1] = sv.y;
        st.rad[2] = sv.z;
      }
    }
    march(r, o, d, stepSize, jitter((uint32_t)px, (uint32_t)py, P.time), end, st);
    result = make_float4(st.rad[0], st.rad[1], st.rad[2], 1.0f - st.transmittance);
    visits = st.visits;
    shades = st.shades;
  }
  out[(size_t)py * P.width + px] = result;
}
__global__ void __launch_bounds__(256) k_march(RenderArgs r, float4 *out, const float4 *slices, int sw, int sh,
                                               unsigned long long *stats) {
  const RenderParams &P = r.P;
  const int px = blockIdx.x * 16 + (threadIdx.x & 15), py = blockIdx.y * 16 + (threadIdx.x >> 4);
  int visits = 0, shades = 0;
  if (px < P.width && py < P.height) marchPixel(r, out, slices, sw, sh, px, py, visits, shades);
  if (stats) countSamples(stats + 2, visits, shades);
}


// ---- scene lights (scene-lights.wgsl:36-106, SceneLights.ts:127-218) ----------------------
// Flames light the scene from anchors (emitters, burning explosions). Every active brick
// samples its fields on a 4^3 lattice (cell centers B/8, 3B/8, ... of each axis), weighs
// each sample by the luminance of its emission, and gives its totals to the anchor nearest
// the brick's center, where height above an anchor counts a quarter.

// anchorDistance (scene-lights.wgsl:19-23).
__device__ __forceinline__ float anchorDistance(const float c[3], const float4 &a) {
  const float dx = c[0] - a.x, dy = c[1] - a.y, dz = c[2] - a.z;
  const float rise = dy > 0.0f ? dy * 0.25f : dy;
  return dx * dx + dz * dz + rise * rise;
}

// One 64-lane workgroup per active brick: its nearest anchor (`brickAnchor`, -1 for a
// brick without a page) and its seven totals: luminance, luminance-weighted offset from the
// anchor, rgb, each summed over the lanes by a pairwise tree.
__global__ void __launch_bounds__(64) k_light_bricks(RenderArgs r, const int *tileRec, const float4 *anchors, int count,
                                                     float *partials, int *brickAnchor) {
  const int bi = (int)blockIdx.x;
  if (bi >= r.counts[0]) return;
  const int lane = (int)threadIdx.x;
  const int id = r.active[bi];
  const int page = r.pages[id];
  if (page < 0) {
    if (lane == 0) brickAnchor[bi] = -1;
    return;
  }
  __shared__ float nearD[64];
  __shared__ int nearI[64];
  __shared__ float4 pos[64];
  __shared__ float3 col[64];
  const RenderParams &P = r.P;
  const int *tr = tileRec + (size_t)(id >> 6) * TILE_RECORD_WORDS;
  const int l = id & 63;
  const int brick[3] = {tr[0] * 4 + (l & 3), tr[1] * 4 + ((l >> 2) & 3), tr[2] * 4 + (l >> 4)};
  const int stride = r.B >> 2;
  const int local[3] = {(lane & 3) * stride + stride / 2, ((lane >> 2) & 3) * stride + stride / 2,
                        (lane >> 4) * stride + stride / 2};
  float4 f = r.F[(size_t)pageToSlot(page, r.side) * r.nF + ((local[2] * r.B + local[1]) * r.B + local[0])];
  if (r.coarse) {
    // The simulation's coarse density sampled at the lattice point (scene-lights.wgsl:54-61),
    // cells of neighbor bricks through the directory.
    const float ratio = (float)(r.B / r.bs);
    const int n = r.bs, mk = n - 1;
    float q[3], fr[3];
    int base[3];
    for (int a = 0; a < 3; a++) {
      q[a] = (float)(brick[a] * n) + ((float)local[a] + 0.5f) / ratio - 0.5f;
      const float g = floorf(q[a]);
      base[a] = (int)g;
      fr[a] = q[a] - g;
    }
    float v = 0.0f;
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          const int c[3] = {base[0] + x, base[1] + y, base[2] + z};
          const int e = entryLookup(r.dir, c[0] >> r.sS, c[1] >> r.sS, c[2] >> r.sS);
          if (e < 0) continue;
          const float w = (x ? fr[0] : 1.0f - fr[0]) * (y ? fr[1] : 1.0f - fr[1]) * (z ? fr[2] : 1.0f - fr[2]);
          v += r.S[(size_t)entrySlot(e, r.side) * r.nS + (((c[2] & mk) * n + (c[1] & mk)) * n + (c[0] & mk))] * w;
        }
    f.x = v;
  }
  float rgb[3];
  emission(r, f, rgb);
  const float luminance = rgb[0] * 0.2126f + rgb[1] * 0.7152f + rgb[2] * 0.0722f;
  const float center[3] = {((float)brick[0] + 0.5f) * P.brickSize, ((float)brick[1] + 0.5f) * P.brickSize,
                           ((float)brick[2] + 0.5f) * P.brickSize};
  float bestD = 1e38f;
  int bestI = 0;
  for (int i = lane; i < count; i += 64) {
    const float d = anchorDistance(center, anchors[i]);
    if (d < bestD) {
      bestD = d;
      bestI = i;
    }
  }
  nearD[lane] = bestD;
  nearI[lane] = bestI;
  __syncthreads();
  for (int reach = 32; reach > 0; reach >>= 1) {
    if (lane < reach) {
      const float od = nearD[lane + reach];
      const int oi = nearI[lane + reach];
      if (od < nearD[lane] || (od == nearD[lane] && oi < nearI[lane])) {
        nearD[lane] = od;
        nearI[lane] = oi;
      }
    }
    __syncthreads();
  }
  const int anchor = nearI[0];
  const float4 a = anchors[anchor];
  const float p[3] = {((float)(brick[0] * r.B + local[0]) + 0.5f) * P.voxel,
                      ((float)(brick[1] * r.B + local[1]) + 0.5f) * P.voxel,
                      ((float)(brick[2] * r.B + local[2]) + 0.5f) * P.voxel};
  pos[lane] = make_float4(luminance, (p[0] - a.x) * luminance, (p[1] - a.y) * luminance, (p[2] - a.z) * luminance);
  col[lane] = make_float3(rgb[0], rgb[1], rgb[2]);
  __syncthreads();
  for (int reach = 32; reach > 0; reach >>= 1) {
    if (lane < reach) {
      const float4 o = pos[lane + reach];
      const float3 oc = col[lane + reach];
      pos[lane] = make_float4(pos[lane].x + o.x, pos[lane].y + o.y, pos[lane].z + o.z, pos[lane].w + o.w);
      col[lane] = make_float3(col[lane].x + oc.x, col[lane].y + oc.y, col[lane].z + oc.z);
    }
    __syncthreads();
  }
  if (lane == 0) {
    float *out = partials + (size_t)bi * 8;
    out[0] = pos[0].x;
    out[1] = pos[0].y;
    out[2] = pos[0].z;
    out[3] = pos[0].w;
    out[4] = col[0].x;
    out[5] = col[0].y;
    out[6] = col[0].z;
    out[7] = 0.0f;
    brickAnchor[bi] = anchor;
  }
}

// Per anchor, the totals of LIGHT_BLOCK consecutive bricks by a pairwise tree (bricks of
// other anchors add zero): blockSums[(block * count + anchor) * 8 + k]. The order is fixed,
// so the result does not depend on scheduling (the reference adds with float atomics).
__global__ void __launch_bounds__(LIGHT_BLOCK) k_light_blocks(const int *counts, const float *partials,
                                                              const int *brickAnchor, int count, float *blockSums) {
  const int n = counts[0];
  const int block = (int)blockIdx.x;
  if (block * LIGHT_BLOCK >= n && !(n == 0 && block == 0)) return;
  __shared__ float red[LIGHT_BLOCK];
  const int lane = (int)threadIdx.x, bi = block * LIGHT_BLOCK + lane;
  const int mine = bi < n ? brickAnchor[bi] : -1;
  for (int a = 0; a < count; a++)
    for (int k = 0; k < 7; k++) {
      red[lane] = mine == a ? partials[(size_t)bi * 8 + k] : 0.0f;
      __syncthreads();
      for (int reach = LIGHT_BLOCK / 2; reach > 0; reach >>= 1) {
        if (lane < reach) red[lane] += red[lane + reach];
        __syncthreads();
      }
      if (lane == 0) blockSums[((size_t)block * count + a) * 8 + k] = red[0];
      __syncthreads();
    }
}

// The anchors' totals over the blocks: lane j adds blocks j, j + 256, ... in order, then a
// pairwise tree over the lanes. sums[anchor * 8 + k], the eighth float zero.
__global__ void __launch_bounds__(LIGHT_BLOCK) k_light_total(const int *counts, const float *blockSums, int count,
                                                             float *sums) {
  __shared__ float red[LIGHT_BLOCK];
  const int lane = (int)threadIdx.x;
  const int blocks = counts[0] > 0 ? (counts[0] + LIGHT_BLOCK - 1) / LIGHT_BLOCK : 0;
  for (int a = 0; a < count; a++) {
    for (int k = 0; k < 7; k++) {
      float acc = 0.0f;
      for (int b = lane; b < blocks; b += LIGHT_BLOCK) acc += blockSums[((size_t)b * count + a) * 8 + k];
      red[lane] = acc;
      __syncthreads();
      for (int reach = LIGHT_BLOCK / 2; reach > 0; reach >>= 1) {
        if (lane < reach) red[lane] += red[lane + reach];
        __syncthreads();
      }
      if (lane == 0) sums[(size_t)a * 8 + k] = red[0];
      __syncthreads();
    }
    if (lane == 0) sums[(size_t)a * 8 + 7] = 0.0f;
  }
}

}  // namespace fp
