// Transport, combustion, forces, divergence and projection over the active bricks
// (fluid-velocity-transport.wgsl, fluid-scalar-transport.wgsl, fluid-field-evolution.wgsl,
// fluid-content.wgsl, fluid-forces.wgsl, force-fields.wgsl). One workgroup covers one
// active brick (or a 256-cell chunk of one); the brick's neighbor record is staged in LDS.
#include "kernels.hpp"

namespace fp {

#define BRICK_PREAMBLE(CHUNKS)                              \
  const int item = xcdItem(vb, s.counts[0] * (CHUNKS));     \
  if (item < 0) return;                                     \
  const int bi = item / (CHUNKS);                           \
  const int chunk = item % (CHUNKS);                        \
  (void)chunk;                                              \
  const int id = s.active[bi];                              \
  __shared__ int rec[HOME_WORDS];                                   \
  Ctx c;                                                    \
  c.dir = s.dir;                                            \
  c.L = s.L;                                                \
  c.side = s.L.side;                                        \
  loadHome(s.dir, id, rec, c.home, s.L.side);               \
  const int slot = entrySlot(rec[13], s.L.side);

__device__ __forceinline__ void velCellCoords(const Ctx &c, int l, int out[3]) {
  const int b = c.L.b, sV = c.L.sV;
  out[0] = c.home.bx * b + (l & (b - 1));
  out[1] = c.home.by * b + ((l >> sV) & (b - 1));
  out[2] = c.home.bz * b + (l >> (2 * sV));
}
__device__ __forceinline__ void fieldCellCoords(const Ctx &c, int l, int out[3]) {
  const int B = c.L.B, sF = c.L.sF;
  out[0] = c.home.bx * B + (l & (B - 1));
  out[1] = c.home.by * B + ((l >> sF) & (B - 1));
  out[2] = c.home.bz * B + (l >> (2 * sF));
}

// ---- velocity transport -------------------------------------------------------------

// advectVelocityComponent (fluid-velocity-transport.wgsl:28-41). Also returns the face
// velocity `fv` and the donor bounds [lo, hi] of the departure cell, which the corrector
// would otherwise recompute from the same backtrace.
__device__ __forceinline__ float advectComponent(const Ctx &c, const VelWindow &w, const Params &P, const float *V,
                                 const int id[3], const float pos[3], int comp, const float orig[3], float fv[3],
                                 float &lo, float &hi) {
  float off[3] = {0.f, 0.f, 0.f};
  off[comp] = 0.5f * P.hv;
  float fp[3], mid[3], v[3];
  for (int a = 0; a < 3; a++) fp[a] = pos[a] + off[a];
  velocityOnFace(c, w, V, id, comp, orig, fv);
  for (int a = 0; a < 3; a++) mid[a] = fp[a] - 0.5f * P.dt * fv[a];
  velocityAt(c, w, P, V, mid[0], mid[1], mid[2], v);
  float sp[3];
  for (int a = 0; a < 3; a++) sp[a] = (fp[a] - P.dt * v[a] - off[a]) * P.invHv;
  return windowVelBounds(c, w, V, comp, sp[0], sp[1], sp[2], lo, hi);
}

__device__ __forceinline__ void advect_velocity_body(const StepArgs &s, const int vb) {
  BRICK_PREAMBLE(1)
  // Every sample reads V: the whole velocity brick's window.
  __shared__ float vwin[3 * VW_CELLS];
  VelWindow w;
  stageVelWindow(c, s.p.V, 0, s.L.B - 1, vwin, w);
  const int nV = s.L.nV;
  float *out = s.p.Vadv + slotBase(slot * 3, s.L.sV);
  const float *vs = s.p.V + slotBase(slot * 3, s.L.sV);
  for (int l = threadIdx.x; l < nV; l += blockDim.x) {
    int cc[3];
    velCellCoords(c, l, cc);
    float pos[3] = {((float)cc[0] + 0.5f) * s.P.hv, ((float)cc[1] + 0.5f) * s.P.hv, ((float)cc[2] + 0.5f) * s.P.hv};
    if (solidAtNear(c, s.P, s.p.solids, pos[0], pos[1], pos[2])) {
      out[l] = out[nV + l] = out[2 * nV + l] = 0.f;
      continue;
    }
    float orig[3] = {vs[l], vs[nV + l], vs[2 * nV + l]};
    float4 *tr = s.p.Vtr + slotBase(slot * 4, s.L.sV) + l;
    float hi[3];
    for (int k = 0; k < 3; k++) {
      float fv[3], lo;
      out[k * nV + l] = advectComponent(c, w, s.P, s.p.V, cc, pos, k, orig, fv, lo, hi[k]);
      tr[k * nV] = make_float4(fv[0], fv[1], fv[2], lo);
    }
    tr[3 * nV] = make_float4(hi[0], hi[1], hi[2], 0.f);
  }
}
FP_KERNEL_PAIR(advect_velocity, 256, StepArgs)


// correctVelocityComponent (fluid-velocity-transport.wgsl:114-156), with the face velocity
// and donor bounds that advection left in the backtrace planes; velocity samples through
// the window `wv`, the advected velocity through `wa`.
__device__ __forceinline__ float correctComponent(const Ctx &c, const VelWindow &wv, const VelWindow &wa,
                                  const Params &P, const Pools &p, const float pos[3],
                                  int comp, const float orig[3], float predicted, const float fv[3], float lo,
                                  float hi) {
  float off[3] = {0.f, 0.f, 0.f};
  off[comp] = 0.5f * P.hv;
  float fp[3];
  for (int a = 0; a < 3; a++) fp[a] = pos[a] + off[a];
  if (lo == hi && predicted == lo) return predicted;
  float fmid[3], fw[3], sp[3];
  for (int a = 0; a < 3; a++) fmid[a] = fp[a] + 0.5f * P.dt * fv[a];
  velocityAt(c, wv, P, p.V, fmid[0], fmid[1], fmid[2], fw);
  for (int a = 0; a < 3; a++) sp[a] = (fp[a] + P.dt * fw[a] - off[a]) * P.invHv;
  float reversed = windowVel(c, wa, p.Vadv, comp, sp[0], sp[1], sp[2]);
  float corrected = predicted + 0.5f * (orig[comp] - reversed);
  return (corrected < lo || corrected > hi) ? predicted : corrected;
}

__device__ __forceinline__ void correct_velocity_body(const StepArgs &s, const int vb) {
  BRICK_PREAMBLE(1)
  const int nV = s.L.nV;
  float *out = s.p.Vcor + slotBase(slot * 3, s.L.sV);
  const float *vs = s.p.V + slotBase(slot * 3, s.L.sV);
  const float *as = s.p.Vadv + slotBase(slot * 3, s.L.sV);
  // The midpoint samples read the velocity through the brick's window; the reversed
  // sample's window of the advected planes is off: a second 9 KB of LDS a workgroup costs
  // more resident wavefronts than it saves (measured).
  __shared__ float vwin[3 * VW_CELLS];
  VelWindow wv, wa;
  stageVelWindow(c, s.p.V, 0, s.L.B - 1, vwin, wv);
  velWindowOff(vwin, wa);
  for (int l = threadIdx.x; l < nV; l += blockDim.x) {
    int cc[3];
    velCellCoords(c, l, cc);
    float pos[3] = {((float)cc[0] + 0.5f) * s.P.hv, ((float)cc[1] + 0.5f) * s.P.hv, ((float)cc[2] + 0.5f) * s.P.hv};
    if (solidAtNear(c, s.P, s.p.solids, pos[0], pos[1], pos[2])) {
      out[l] = out[nV + l] = out[2 * nV + l] = 0.f;
      continue;
    }
    float orig[3] = {vs[l], vs[nV + l], vs[2 * nV + l]};
    const float4 *tr = s.p.Vtr + slotBase(slot * 4, s.L.sV) + l;
    const float4 hi = tr[3 * nV];
    const float his[3] = {hi.x, hi.y, hi.z};
    for (int k = 0; k < 3; k++) {
      const float4 t = tr[k * nV];
      const float fv[3] = {t.x, t.y, t.z};
      out[k * nV + l] = correctComponent(c, wv, wa, s.P, s.p, pos, k, orig, as[k * nV + l], fv, t.w, his[k]);
    }
  }
}
FP_KERNEL_PAIR(correct_velocity, 256, StepArgs)


// ---- scalar transport -----------------------------------------------------------------

// contentSupport (fluid-content.wgsl:29-39) on a grid of bricks `side` cells a side,
// velocity cells `ratio` of its cells wide and cells 1 / invH wide.
__device__ __forceinline__ uint32_t supportOn(int side, float ratio, float invH, float dt, const int id[3],
                                              const float vel[3]) {
  float sidef = (float)side;
  bool low[3], high[3];
  for (int a = 0; a < 3; a++) {
    float local = (float)(id[a] & (side - 1)) + 0.5f;
    float pad = fminv(sidef, fmaxv(2.0f, 2.0f * ratio) + fabsv(vel[a]) * dt * invH);
    low[a] = local < pad;
    high[a] = local + pad >= sidef;
  }
  uint32_t row = 2u | (low[0] ? 1u : 0u) | (high[0] ? 4u : 0u);
  uint32_t plane = (row << 3) | (low[1] ? row : 0u) | (high[1] ? row << 6 : 0u);
  return (plane << 9) | (low[2] ? plane : 0u) | (high[2] ? plane << 18 : 0u);
}
__device__ __forceinline__ uint32_t contentSupport(const Layout &L, const Params &P, const int id[3], const float vel[3]) {
  return supportOn(L.B, P.hv / P.hf, P.invHf, P.dt, id, vel);
}


struct Donors {
  float mn[3], mx[3], lmn, lmx;  // smoke, heat, fuel; lifetime
};

// sameDonors (fluid-scalar-transport.wgsl:64-67).
__device__ __forceinline__ bool uniformDonors(const Donors &dn) {
  return dn.mn[0] == dn.mx[0] && dn.mn[1] == dn.mx[1] && dn.mn[2] == dn.mx[2] && dn.lmn == dn.lmx;
}
// predictScalarsWithBounds (fluid-scalar-transport.wgsl:38-62).
__device__ __forceinline__ float4 predict(const Ctx &c, const float4 *F, const int base[3], const float frac[3], Donors &dn) {
  for (int k = 0; k < 3; k++) {
    dn.mn[k] = 65504.0f;
    dn.mx[k] = -65504.0f;
  }
  dn.lmn = 1.0f;
  dn.lmx = 0.0f;
  float4 value = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 cv[8];
  fieldCorners(c, F, base[0], base[1], base[2], cv);
  {
    float ch[4][8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      ch[0][k] = cv[k].x;
      ch[1][k] = cv[k].y;
      ch[2][k] = cv[k].w;
      ch[3][k] = cv[k].z;
    }
    range8(ch[0], dn.mn[0], dn.mx[0]);
    range8(ch[1], dn.mn[1], dn.mx[1]);
    range8(ch[2], dn.mn[2], dn.mx[2]);
    range8(ch[3], dn.lmn, dn.lmx);
  }
  for (int z = 0; z < 2; z++)
    for (int y = 0; y < 2; y++)
      for (int x = 0; x < 2; x++) {
        float4 smp = cv[(z * 2 + y) * 2 + x];
        float wx = x ? frac[0] : 1.0f - frac[0], wy = y ? frac[1] : 1.0f - frac[1], wz = z ? frac[2] : 1.0f - frac[2];
        value.x += smp.x * wx * wy * wz;
        value.y += smp.y * wx * wy * wz;
        value.z += smp.z * wx * wy * wz;
        value.w += smp.w * wx * wy * wz;
      }
  return value;
}

__device__ __forceinline__ void traceBack(const Ctx &c, const Params &P, const float *V, const float pos[3],
                                          const float cv[3], float out[3]) {
  float m[3], v[3];
  for (int a = 0; a < 3; a++) m[a] = pos[a] - 0.5f * P.dt * cv[a];
  velocityAt(c, P, V, m[0], m[1], m[2], v);
  for (int a = 0; a < 3; a++) out[a] = pos[a] - P.dt * v[a];
}
__device__ __forceinline__ void traceForward(const Ctx &c, const Params &P, const float *V, const float pos[3],
                                             const float cv[3], float out[3]) {
  float m[3], v[3];
  for (int a = 0; a < 3; a++) m[a] = pos[a] + 0.5f * P.dt * cv[a];
  velocityAt(c, P, V, m[0], m[1], m[2], v);
  for (int a = 0; a < 3; a++) out[a] = pos[a] + P.dt * v[a];
}
// The forward trace through a workgroup's velocity window.
__device__ __forceinline__ void traceForward(const Ctx &c, const VelWindow &w, const Params &P, const float *V,
                                             const float pos[3], const float cv[3], float out[3]) {
  float m[3], v[3];
  for (int a = 0; a < 3; a++) m[a] = pos[a] + 0.5f * P.dt * cv[a];
  velocityAt(c, w, P, V, m[0], m[1], m[2], v);
  for (int a = 0; a < 3; a++) out[a] = pos[a] + P.dt * v[a];
}
// The backward trace through a workgroup's velocity window.
__device__ __forceinline__ void traceBack(const Ctx &c, const VelWindow &w, const Params &P, const float *V,
                                          const float pos[3], const float cv[3], float out[3]) {
  float m[3], v[3];
  for (int a = 0; a < 3; a++) m[a] = pos[a] - 0.5f * P.dt * cv[a];
  velocityAt(c, w, P, V, m[0], m[1], m[2], v);
  for (int a = 0; a < 3; a++) out[a] = pos[a] - P.dt * v[a];
}
__device__ __forceinline__ void departureCell(const Params &P, const int id[3], const float pos[3],
                                              const float dep[3], int base[3], float frac[3]) {
  for (int a = 0; a < 3; a++) {
    float cell = (float)id[a] + (dep[a] - pos[a]) * P.invHf;
    float fl = floorf(cell);
    base[a] = (int)fl;
    frac[a] = cell - fl;
  }
}

// The field-cell kernels give each brick max(1, nF / n) chunks of n lanes (the workgroup's
// size), one cell a lane: chunk k holds the brick's cells [n k, n k + n), a slab of whole z
// layers (or part of one), whose local z range this is.
__device__ __forceinline__ void chunkSlab(const Layout &L, int chunk, int &zlo, int &zhi) {
  const int n = (int)blockDim.x;
  const int first = chunk * n, last = (first + n < L.nF ? first + n : L.nF) - 1;
  zlo = first >> (2 * L.sF);
  zhi = last >> (2 * L.sF);
}

__device__ __forceinline__ void advect_scalars_body(const StepArgs &s, const int vb, int chunks) {
  BRICK_PREAMBLE(chunks)
  (void)chunks;
  const int nF = s.L.nF;
  __shared__ float vwin[3 * VW_CELLS];
  VelWindow w;
  int zlo, zhi;
  chunkSlab(s.L, chunk, zlo, zhi);
  stageVelWindow(c, s.p.V, zlo, zhi, vwin, w);
  const int l = chunk * (int)blockDim.x + threadIdx.x;
  if (l >= nF) return;
  int cc[3];
  fieldCellCoords(c, l, cc);
  float pos[3] = {((float)cc[0] + 0.5f) * s.P.hf, ((float)cc[1] + 0.5f) * s.P.hf, ((float)cc[2] + 0.5f) * s.P.hf};
  float cv[3], dep[3];
  velocityAt(c, w, s.P, s.p.V, pos[0], pos[1], pos[2], cv);
  traceBack(c, w, s.P, s.p.V, pos, cv, dep);
  float4 *out = s.p.Fp + slotBase(slot, s.L.sF) + l;
  if (!s.P.maccormack) {
    float4 v = sampleField(c, s.p.F, dep[0] * s.P.invHf, dep[1] * s.P.invHf, dep[2] * s.P.invHf);
    *out = make_float4(s.P.coarse ? 0.f : v.x, v.y, clampv(v.z, 0.f, 1.f), v.w);
    return;
  }
  int base[3];
  float frac[3];
  departureCell(s.P, cc, pos, dep, base, frac);
  // The content-support mask, the corrector's only other use of the cell velocity.
  const uint32_t support = contentSupport(s.L, s.P, cc, cv);
  Donors dn;
  float4 v = predict(c, s.p.F, base, frac, dn);
  // Coarse smoke: the fine X channel only stages the step's smoke production
  // (fluid-scalar-transport.wgsl:127, :137, :191-199).
  *out = make_float4(s.P.coarse ? 0.f : v.x, v.y, s.P.lifetimeZero ? 0.f : v.z, v.w);
  // Stores in the order that ends the bounds' registers before the forward trace (79
  // VGPRs, six wavefronts a SIMD, against 87 and five with the trace before the stores).
  float4 *don = s.p.Don + slotBase(slot * 3, s.L.sF) + l;
  don[nF] = make_float4(dn.mn[0], dn.mn[1], dn.mn[2], dn.lmx);
  don[2 * nF] = make_float4(dn.mx[0], dn.mx[1], dn.mx[2], __uint_as_float(support));
  // The corrector's forward trace (correctedFineScalars, fluid-scalar-transport.wgsl:150-152)
  // depends only on the velocity, so it runs here through the window; the corrector needs
  // it only where the donors differ.
  const float lmn = dn.lmn;
  float fw[3] = {0.f, 0.f, 0.f};
  if (!uniformDonors(dn)) traceForward(c, w, s.P, s.p.V, pos, cv, fw);
  don[0] = make_float4(fw[0], fw[1], fw[2], lmn);
}
FP_KERNEL_PAIR1(advect_scalars, 256, StepArgs, int, chunks)


// correctedFineScalars (fluid-scalar-transport.wgsl:144-188), with the donor bounds and the
// forward departure point `fw` that advection left in the backtrace planes.
__device__ __forceinline__ float4 correctedScalars(const Ctx &c, const Params &P, const Pools &p, const int id[3],
                                   int slot, int l, const float pos[3], const float fw[3], const Donors &dn) {
  const float4 adv = p.Fp[slotBase(slot, c.L.sF) + l];
  if (!P.maccormack) return adv;
  if (uniformDonors(dn)) return make_float4(dn.mn[0], dn.mn[1], clampv(dn.lmn, 0.f, 1.f), dn.mn[2]);
  const float4 orig = p.F[slotBase(slot, c.L.sF) + l];
  float lifetime = adv.z;
  float rev[3];
  if (dn.lmn != dn.lmx || adv.z != dn.lmn) {
    // sampleScalarsExplicit (fluid-scalar-transport.wgsl:6-21) at the forward cell.
    int bs[3];
    float fr[3];
    for (int a = 0; a < 3; a++) {
      float cell = (float)id[a] + (fw[a] - pos[a]) * P.invHf;
      float fl = floorf(cell);
      bs[a] = (int)fl;
      fr[a] = cell - fl;
    }
    float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 cvs[8];
    fieldCorners(c, p.Fp, bs[0], bs[1], bs[2], cvs);
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          float4 smp = cvs[(z * 2 + y) * 2 + x];
          float wx = x ? fr[0] : 1.0f - fr[0], wy = y ? fr[1] : 1.0f - fr[1], wz = z ? fr[2] : 1.0f - fr[2];
          val.x += smp.x * wx * wy * wz;
          val.y += smp.y * wx * wy * wz;
          val.z += smp.z * wx * wy * wz;
          val.w += smp.w * wx * wy * wz;
        }
    rev[0] = val.x;
    rev[1] = val.y;
    rev[2] = val.w;
    float cl = adv.z + 0.5f * (orig.z - val.z);
    lifetime = (cl < dn.lmn || cl > dn.lmx) ? adv.z : cl;
  } else {
    float4 val = sampleField(c, p.Fp, fw[0] * P.invHf, fw[1] * P.invHf, fw[2] * P.invHf);
    rev[0] = val.x;
    rev[1] = val.y;
    rev[2] = val.w;
  }
  const float a0 = adv.x, a1 = adv.y, a3 = adv.w;
  float c0 = a0 + 0.5f * (orig.x - rev[0]);
  float c1 = a1 + 0.5f * (orig.y - rev[1]);
  float c3 = a3 + 0.5f * (orig.w - rev[2]);
  float b0 = (c0 < dn.mn[0] || c0 > dn.mx[0]) ? a0 : c0;
  float b1 = (c1 < dn.mn[1] || c1 > dn.mx[1]) ? a1 : c1;
  float b3 = (c3 < dn.mn[2] || c3 > dn.mx[2]) ? a3 : c3;
  return make_float4(b0, b1, clampv(lifetime, 0.f, 1.f), b3);
}

// ---- emission and combustion (fluid-sources.wgsl, flame-common.wgsl) -------------------

__device__ __forceinline__ float emitterWeight(const Emitter &e, const float w[3]) {
  // Past the padded size on any axis |d| >= 1 there, so r >= 1: no weight.
  for (int a = 0; a < 3; a++)
    if (fabsv(w[a] - e.pos[a]) >= fmaxv(e.size[a], 0.03f)) return 0.f;
  float d[3];
  for (int a = 0; a < 3; a++) d[a] = (w[a] - e.pos[a]) * e.invSize[a];
  float r = e.box ? fmaxv(fabsv(d[0]), fmaxv(fabsv(d[1]), fabsv(d[2]))) : length3(d[0], d[1], d[2]);
  if (r >= 1.0f) return 0.f;
  return 1.0f - smoothstepv(0.5f, 1.0f, r);
}

__device__ __forceinline__ float burstGradient(uint32_t seedBits, int cx, int cy, int cz, float ox, float oy, float oz) {
  uint32_t h = (uint32_t)cx * 374761393u + (uint32_t)cy * 668265263u + (uint32_t)cz * 2246822519u +
               seedBits * 3266489917u;
  h = (h ^ (h >> 13)) * 1274126177u;
  h = (h ^ (h >> 16)) & 15u;
  float a = h < 8u ? ox : oy;
  float b = h < 4u ? oy : ((h == 12u || h == 14u) ? ox : oz);
  return ((h & 1u) == 0u ? a : -a) + ((h & 2u) == 0u ? b : -b);
}
__device__ __forceinline__ float fade(float f) { return f * f * f * (f * (f * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float burstPerlin(uint32_t seedBits, float px, float py, float pz) {
  float flx = floorf(px), fly = floorf(py), flz = floorf(pz);
  int cx = (int)flx, cy = (int)fly, cz = (int)flz;
  float fx = px - flx, fy = py - fly, fz = pz - flz;
  float sx = fade(fx), sy = fade(fy), sz = fade(fz);
  float a = burstGradient(seedBits, cx, cy, cz, fx, fy, fz);
  float b = burstGradient(seedBits, cx + 1, cy, cz, fx - 1.0f, fy, fz);
  float cc = burstGradient(seedBits, cx, cy + 1, cz, fx, fy - 1.0f, fz);
  float d = burstGradient(seedBits, cx + 1, cy + 1, cz, fx - 1.0f, fy - 1.0f, fz);
  float e = burstGradient(seedBits, cx, cy, cz + 1, fx, fy, fz - 1.0f);
  float g = burstGradient(seedBits, cx + 1, cy, cz + 1, fx - 1.0f, fy, fz - 1.0f);
  float h = burstGradient(seedBits, cx, cy + 1, cz + 1, fx, fy - 1.0f, fz - 1.0f);
  float j = burstGradient(seedBits, cx + 1, cy + 1, cz + 1, fx - 1.0f, fy - 1.0f, fz - 1.0f);
  return mixv(mixv(mixv(a, b, sx), mixv(cc, d, sx), sy), mixv(mixv(e, g, sx), mixv(h, j, sx), sy), sz);
}
__device__ __forceinline__ float burstPattern(uint32_t seedBits, float px, float py, float pz) {
  float n = (burstPerlin(seedBits, px, py, pz) +
             0.35f * burstPerlin(seedBits, px * 2.0f + 5.2f, py * 2.0f + 9.7f, pz * 2.0f + 3.1f)) /
            1.35f;
  return clampv(0.5f + 0.5f * n, 0.f, 1.f);
}
__device__ __forceinline__ void burstVariation(uint32_t seedBits, const Emitter &e, const float w[3], float out[2]) {
  if (e.strength == 0.f) {
    out[0] = out[1] = 1.f;
    return;
  }
  const float off[3] = {3.17f, 7.53f, 11.29f};
  float q[3];
  for (int a = 0; a < 3; a++) q[a] = (w[a] - e.pos[a]) * e.invPeriod + off[a];
  float s0 = burstPattern(seedBits, q[0], q[1], q[2]);
  float s1 = burstPattern(seedBits, q[0] + 17.3f, q[1] + 2.8f, q[2] + 8.6f);
  out[0] = 1.0f - e.strength * s0;
  out[1] = 1.0f - e.strength * s1;
}

// Flame ramps staged in LDS: knots [channel][knot][x, y], then the knot counts.
struct Ramps {
  const float *knots;
  const int *counts;
};
__device__ __forceinline__ float rampIntegral(const Ramps &R, int ch, float lo, float hi) {
  float area = 0.f;
  const float *k = R.knots + ch * 16;
  for (int i = 1; i < R.counts[ch]; i++) {
    float ax = k[2 * (i - 1)], ay = k[2 * (i - 1) + 1];
    float bx = k[2 * i], by = k[2 * i + 1];
    float left = fmaxv(lo, ax), right = fminv(hi, bx);
    if (right > left) {
      float y0 = mixv(ay, by, (left - ax) / (bx - ax));
      float y1 = mixv(ay, by, (right - ax) / (bx - ax));
      area += (right - left) * (y0 + y1) * 0.5f;
    }
  }
  return area;
}
__device__ __forceinline__ void evaluateFlame(const Params &P, const Ramps &R, float before, float &after,
                                              float prod[3]) {
  if (before <= 0.f) {
    after = 0.f;
    prod[0] = prod[1] = prod[2] = 0.f;
    return;
  }
  after = fmaxv(0.f, before - P.lifeStep);
  prod[0] = P.fsmoke * P.lifespan * rampIntegral(R, 0, after, before);
  prod[1] = P.fheat * P.lifespan * rampIntegral(R, 1, after, before);
  prod[2] = P.fexpansion * (P.lifespan * rampIntegral(R, 2, after, before) / P.fdt);
}

// meshSourceAt (fluid-sources.wgsl:8-22): the index of the mesh emission record of the field
// cell at world position w (fieldCell, solids.wgsl:9-11), 0 (the zero record) for none.
__device__ __forceinline__ int meshRecordAt(const StepArgs &s, const float w[3]) {
  const int x = (int)floorf(w[0] * s.P.invHf + 0.25f), y = (int)floorf(w[1] * s.P.invHf + 0.25f),
            z = (int)floorf(w[2] * s.P.invHf + 0.25f);
  const int sh = s.L.sF, m = s.L.B - 1;
  const int id = brickIdOf(s.dir, x >> sh, y >> sh, z >> sh);
  if (id < 0) return 0;
  const int table = s.meshTables[id];
  if (table < 0) return 0;
  return s.meshTables[s.brickCount + table * s.L.nF + ((z & m) * s.L.B + (y & m)) * s.L.B + (x & m)];
}

// evolveFields (fluid-field-evolution.wgsl:9-84) with content recording.
__device__ __forceinline__ void evolve_fields_body(const StepArgs &s, const int vb, int chunks) {
  BRICK_PREAMBLE(chunks)
  __shared__ uint32_t busy, expands;
  if (threadIdx.x == 0) {
    busy = 0u;
    expands = 0u;
  }
  __syncthreads();
  const Params &P = s.P;
  const int nF = s.L.nF;
  __shared__ float knots[48];
  __shared__ int knotCounts[3];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
      knotCounts[ch] = P.knotCount[ch];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        knots[ch * 16 + 2 * i] = P.knots[ch][i][0];
        knots[ch * 16 + 2 * i + 1] = P.knots[ch][i][1];
      }
    }
  }
  __syncthreads();
  const Ramps ramps{knots, knotCounts};
  uint32_t myBusy = 0u, myExpands = 0u;
  const float burnKeep = P.burnKeep, lossS = P.lossS, lossH = P.lossH, halfS = P.halfS, halfH = P.halfH;
  // One cell a lane (chunkSlab). Only the lanes whose donors differ trace forward, through
  // the pools: a velocity window staged for the whole workgroup costs more than it saves
  // (measured). The single-simulation step launches 64 lanes a workgroup, a batch 256.
  (void)chunks;
  const int l = chunk * (int)blockDim.x + threadIdx.x;
  const bool live = l < nF;
  // Emitters whose padded box misses every cell center of this slab weigh nothing in it
  // (emitterWeight's first test, decided here for all the slab's cells at once: w - pos is
  // monotone in w): a mask of the others among the first 32; later ones are always tested.
  uint32_t nearby = 0xffffffffu;
  {
    int zlo, zhi;
    chunkSlab(s.L, chunk, zlo, zhi);
    const int B = s.L.B;
    const int lo[3] = {c.home.bx * B, c.home.by * B, c.home.bz * B + zlo};
    const int hi[3] = {lo[0] + B - 1, lo[1] + B - 1, c.home.bz * B + zhi};
    const int count = P.emitterCount < 32 ? P.emitterCount : 32;
    for (int i = 0; i < count; i++) {
      const Emitter &e = s.emitters[i];
      for (int a = 0; a < 3; a++) {
        const float pad = fmaxv(e.size[a], 0.03f);
        const float wlo = ((float)lo[a] + 0.5f) * P.hf, whi = ((float)hi[a] + 0.5f) * P.hf;
        if (wlo - e.pos[a] >= pad || whi - e.pos[a] <= -pad) {
          nearby &= ~(1u << i);
          break;
        }
      }
    }
  }
  float cv[3] = {0.f, 0.f, 0.f}, fw[3] = {0.f, 0.f, 0.f};
  uint32_t support = 0u;
  Donors dn;
  if (live) {
    if (P.maccormack) {
      // The forward departure point, donor bounds and content support from advection.
      const float4 *don = s.p.Don + slotBase(slot * 3, s.L.sF) + l;
      const float4 d0 = don[0], d1 = don[nF], d2 = don[2 * nF];
      fw[0] = d0.x;
      fw[1] = d0.y;
      fw[2] = d0.z;
      support = __float_as_uint(d2.w);
      dn.lmn = d0.w;
      dn.mn[0] = d1.x;
      dn.mn[1] = d1.y;
      dn.mn[2] = d1.z;
      dn.lmx = d1.w;
      dn.mx[0] = d2.x;
      dn.mx[1] = d2.y;
      dn.mx[2] = d2.z;
    }
  }
  if (live) {
    int cc[3];
    fieldCellCoords(c, l, cc);
    float pos[3] = {((float)cc[0] + 0.5f) * P.hf, ((float)cc[1] + 0.5f) * P.hf, ((float)cc[2] + 0.5f) * P.hf};
    float value[4] = {0.f, 0.f, 0.f, 0.f};
    float expansion = 0.f;
    bool injected = false;
    if (!P.maccormack) velocityAt(c, P, s.p.V, pos[0], pos[1], pos[2], cv);
    if (!solidAtNear(c, P, s.p.solids, pos[0], pos[1], pos[2])) {
      float4 tr = correctedScalars(c, P, s.p, cc, slot, l, pos, fw, dn);
      if (P.coarse) tr.x = 0.f;
      float lifetime = tr.z;
      float sp0 = 0.f, sp1 = 0.f, injectedFuel = 0.f;
      for (int i = 0; i < P.emitterCount; i++) {
        if (i < 32 && !((nearby >> i) & 1u)) continue;
        const Emitter &e = s.emitters[i];
        float wgt = emitterWeight(e, pos);
        if (wgt == 0.f) continue;
        float var[2];
        burstVariation(P.seedBits, e, pos, var);
        lifetime = fmaxv(lifetime, e.flame * wgt * var[1]);
        sp0 += e.smoke * wgt * P.dt;
        sp1 += e.heat * var[0] * wgt * P.dt;
        injectedFuel += e.fuel * wgt * P.dt;
      }
      if (P.meshRecords) {
        const float4 mesh = s.meshRecords[2 * meshRecordAt(s, pos)];
        lifetime = fmaxv(lifetime, mesh.x);
        sp0 += mesh.z * P.dt;
        sp1 += mesh.y * P.dt;
        injectedFuel += mesh.w * P.dt;
      }
      float remaining = 0.f, burned = 0.f;
      if (P.fuelEnabled) {
        remaining = tr.w + injectedFuel;
        if (tr.y + sp1 >= P.ignition) {
          lifetime = fmaxv(lifetime, fminv(remaining, 1.0f));
          float unburned = remaining * burnKeep;
          burned = remaining - unburned;
          remaining = unburned;
        }
      }
      float after, prod[3];
      evaluateFlame(P, ramps, clampv(lifetime, 0.f, 1.f), after, prod);
      expansion = prod[2] + P.fexpansion * burned / P.dt;
      float dens = tr.x * lossS + (sp0 + prod[0]) * halfS;
      float heat = tr.y * lossH + (sp1 + prod[1]) * halfH;
      value[0] = fmaxv(dens, 0.f);
      value[1] = fmaxv(heat, 0.f);
      value[2] = after;
      value[3] = remaining;
      injected = sp0 > 0.f || sp1 > 0.f || lifetime > tr.z || injectedFuel > 0.f;
    }
    if (!injected && value[2] == 0.f && value[0] < 0.001f && value[1] < 0.001f && value[3] < 0.001f)
      value[0] = value[1] = value[2] = value[3] = 0.f;
    s.p.Fn[slotBase(slot, s.L.sF) + l] = make_float4(value[0], value[1], value[2], value[3]);
    s.p.E[slotBase(slot, s.L.sF) + l] = expansion;
    if (expansion != 0.f) myExpands = 1u;
    if (fmaxv(fmaxv(P.coarse ? 0.f : value[0], value[2]), value[3]) > P.cutoff)
      myBusy |= P.maccormack ? support : contentSupport(s.L, P, cc, cv);
  }
  // One LDS atomic per wavefront: the lanes' masks are OR-reduced across it first.
  myBusy = waveOr(myBusy);
  myExpands = waveOr(myExpands);
  if ((threadIdx.x & 63) == 0) {
    if (myBusy) atomicOr(&busy, myBusy);
    if (myExpands) atomicOr(&expands, 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t flags = 0u;
    if (busy) {
      flags = CONTENT | (busy << 3);
      uint32_t local = (uint32_t)id & 63u;
      uint32_t bits = (1u << (local & 3u)) | (16u << ((local >> 2) & 3u)) | (256u << (local >> 4));
      atomicOr(&s.activity[s.brickCount + (id >> 6)], bits);
    }
    if (expands) flags |= EXPANDS;
    if (flags) atomicOr(&s.activity[id], flags);
  }
}
// Six wavefronts a SIMD (80 VGPRs, two dwords spilled) against five at the 88 VGPRs the
// compiler would take: 0.235 -> 0.222 ms at the four-fires state.
FP_KERNEL_PAIR1_OCC(evolve_fields, 256, 6, StepArgs, int, chunks)



// ---- forces ---------------------------------------------------------------------------

__device__ __forceinline__ float curlCentered(const Ctx &c, const float *V, int x, int y, int z, int comp) {
  int bx = x - (comp == 0), by = y - (comp == 1), bz = z - (comp == 2);
  float a = velCellNear(c, V, comp, x, y, z), b = velCellNear(c, V, comp, bx, by, bz);
  return 0.5f * (a + b);
}

__device__ __forceinline__ void curl_body(const StepArgs &s, const int vb) {
  if (s.P.vorticity == 0.f) return;
  BRICK_PREAMBLE(1)
  const int nV = s.L.nV;
  const float *V = s.p.Vcor;
  for (int l = threadIdx.x; l < nV; l += blockDim.x) {
    int p[3];
    velCellCoords(c, l, p);
    float hx = 0.5f / s.P.hv;
    int x = p[0], y = p[1], z = p[2];
    float dx0 = (curlCentered(c, V, x + 1, y, z, 1) - curlCentered(c, V, x - 1, y, z, 1)) * hx;
    float dx1 = (curlCentered(c, V, x + 1, y, z, 2) - curlCentered(c, V, x - 1, y, z, 2)) * hx;
    float dy0 = (curlCentered(c, V, x, y + 1, z, 0) - curlCentered(c, V, x, y - 1, z, 0)) * hx;
    float dy1 = (curlCentered(c, V, x, y + 1, z, 2) - curlCentered(c, V, x, y - 1, z, 2)) * hx;
    float dz0 = (curlCentered(c, V, x, y, z + 1, 0) - curlCentered(c, V, x, y, z - 1, 0)) * hx;
    float dz1 = (curlCentered(c, V, x, y, z + 1, 1) - curlCentered(c, V, x, y, z - 1, 1)) * hx;
    float c0 = dy1 - dz1, c1 = dz0 - dx1, c2 = dx0 - dy0;
    s.p.Curl[slotBase(slot, s.L.sV) + l] = make_float4(c0, c1, c2, length3(c0, c1, c2));
  }
}
FP_KERNEL_PAIR(curl, 256, StepArgs)
__global__ void __launch_bounds__(256) k_curl_readback(StepArgs s, int curlBlocks, const int *counts, int *poolState,
                                                       const uint32_t *tileFlags, int cap, int *out) {
  if ((int)blockIdx.x < curlBlocks) {
    curl_body(s, (int)blockIdx.x);
    return;
  }
  readbackLane(counts, poolState, tileFlags, cap, out, ((int)blockIdx.x - curlBlocks) * (int)blockDim.x + (int)threadIdx.x);
}


// curlNoise (fluid-forces.wgsl:3-9): the periodic table, trilinear with wrapping.
__device__ __forceinline__ void curlNoise(const float4 *noise, const float p[3], float out[3]) {
  const float c0[3] = {0.36f, -0.8f, 0.48f}, c1[3] = {0.48f, 0.6f, 0.64f}, c2[3] = {-0.8f, 0.0f, 0.6f};
  float u[3];
  for (int a = 0; a < 3; a++) u[a] = (c0[a] * p[0] + c1[a] * p[1] + c2[a] * p[2]) / 16.0f;
  float f[3];
  int i0[3];
  for (int a = 0; a < 3; a++) {
    float t = u[a] * 64.0f - 0.5f;
    float fl = floorf(t);
    i0[a] = (int)fl;
    f[a] = t - fl;
  }
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  for (int z = 0; z < 2; z++)
    for (int y = 0; y < 2; y++)
      for (int x = 0; x < 2; x++) {
        float w = (x ? f[0] : 1.0f - f[0]) * (y ? f[1] : 1.0f - f[1]) * (z ? f[2] : 1.0f - f[2]);
        int ix = (i0[0] + x) & 63, iy = (i0[1] + y) & 63, iz = (i0[2] + z) & 63;
        float4 t = noise[((size_t)iz * 64 + iy) * 64 + ix];
        s0 += t.x * w;
        s1 += t.y * w;
        s2 += t.z * w;
      }
  out[0] = c0[0] * s0 + c0[1] * s1 + c0[2] * s2;
  out[1] = c1[0] * s0 + c1[1] * s1 + c1[2] * s2;
  out[2] = c2[0] * s0 + c2[1] * s1 + c2[2] * s2;
}

// coverage + externalForces (force-fields.wgsl:50-103).
// meshCovers (force-fields.wgsl:27-47): whether force `index`'s mesh targets cover velocity
// cell `cell`.
__device__ __forceinline__ bool meshCovers(const StepArgs &s, int index, const int cell[3]) {
  const int sh = s.L.sV, side = s.L.b, m = side - 1;
  const int id = brickIdOf(s.dir, cell[0] >> sh, cell[1] >> sh, cell[2] >> sh);
  if (id < 0) return false;
  const uint32_t first = s.forceCoverage[id];
  if (first == 0xffffffffu) return false;
  const uint32_t bit = (uint32_t)(((cell[2] & m) * side + (cell[1] & m)) * side + (cell[0] & m));
  const uint32_t words = (uint32_t)(side * side * side + 31) / 32u;
  for (uint32_t e = 0; e < s.forceCoverage[first]; e++) {
    const uint32_t row = first + 1u + e * (1u + words);
    if (s.forceCoverage[row] == (uint32_t)index) return ((s.forceCoverage[row + 1u + (bit >> 5)] >> (bit & 31u)) & 1u) != 0u;
  }
  return false;
}

__device__ __forceinline__ void externalForces(const StepArgs &s, const int cell[3], const float p[3], float out[3]) {
  const Params &P = s.P;
  out[0] = out[1] = out[2] = 0.f;
  for (int fi = 0; fi < P.forceCount; fi++) {
    const ForceRec &f = s.forces[fi];
    float weight = 1.f;
    if (f.targeted && f.meshTargets && meshCovers(s, fi, cell)) {
      weight = 1.f;
    } else if (f.targeted) {
      float covered = 0.f;
      for (int r = 0; r < f.regionCount; r++) {
        const Region &g = s.regions[f.firstRegion + r];
        float dd[3];
        for (int a = 0; a < 3; a++) dd[a] = fabsv((p[a] - g.center[a]) * g.invSize[a]);
        float dist = g.kind == 3.0f ? fmaxv(dd[0], fmaxv(dd[1], dd[2])) : length3(dd[0], dd[1], dd[2]);
        covered = fmaxv(covered, 1.0f - smoothstepv(0.5f, 1.0f, dist));
      }
      weight = covered;
    }
    if (weight == 0.f) continue;
    float acc[3] = {f.axis[0] * f.strength, f.axis[1] * f.strength, f.axis[2] * f.strength};
    if (f.kind == 1.0f) {
      float q[3] = {p[0] * f.scale * 0.65f + P.seed, p[1] * f.scale * 0.65f + -P.time * 0.7f,
                    p[2] * f.scale * 0.65f + P.seed * 0.25f};
      float n1[3], n2[3];
      curlNoise(s.p.noise, q, n1);
      float q2[3] = {q[0] * 2.03f + 13.7f, q[1] * 2.03f + 5.3f, q[2] * 2.03f + 19.1f};
      curlNoise(s.p.noise, q2, n2);
      for (int a = 0; a < 3; a++) acc[a] = (n1[a] + 0.45f * n2[a]) * f.strength;
    } else if (f.kind >= 2.0f) {
      float delta[3] = {p[0] - f.center[0], p[1] - f.center[1], p[2] - f.center[2]};
      float radial[3] = {delta[0], delta[1], delta[2]};
      if (f.kind == 2.0f) {
        float dp = delta[0] * f.axis[0] + delta[1] * f.axis[1] + delta[2] * f.axis[2];
        for (int a = 0; a < 3; a++) radial[a] = radial[a] - f.axis[a] * dp;
      }
      float distance = length3(radial[0], radial[1], radial[2]);
      float radius = fmaxv(f.scale, 0.01f);
      float dir[3];
      for (int a = 0; a < 3; a++) dir[a] = radial[a] / fmaxv(distance, radius * 0.25f);
      float rr = radial[0] * radial[0] + radial[1] * radial[1] + radial[2] * radial[2];
      float falloff = 1.0f / (1.0f + rr / (radius * radius));
      if (f.kind == 2.0f) {
        float cr[3] = {f.axis[1] * dir[2] - f.axis[2] * dir[1], f.axis[2] * dir[0] - f.axis[0] * dir[2],
                       f.axis[0] * dir[1] - f.axis[1] * dir[0]};
        for (int a = 0; a < 3; a++) acc[a] = (cr[a] * f.strength - dir[a] * f.inward + f.axis[a] * f.lift) * falloff;
      } else {
        for (int a = 0; a < 3; a++) acc[a] = dir[a] * f.strength * falloff;
      }
    }
    for (int a = 0; a < 3; a++) out[a] += acc[a] * weight;
  }
}

// The box average of the field cells under cell `id` of a grid `d` field cells wide: one
// trilinear sample at the shared corner of each 2^3 block (velocityCellAverage,
// fluid-sampling.wgsl:270-286; smokeProduction, smoke.wgsl:24-37).
__device__ __forceinline__ float4 velocityCellAverage(const Ctx &c, const float4 *F, int d, const int id[3]) {
  if (d == 1) return fieldCellNear(c, F, id[0], id[1], id[2]);
  int blocks = d / 2;
  float ratio = (float)d;
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int z = 0; z < blocks; z++)
    for (int y = 0; y < blocks; y++)
      for (int x = 0; x < blocks; x++) {
        float4 v = sampleFieldNear(c, F, (float)id[0] * ratio + 1.0f + 2.0f * (float)x,
                               (float)id[1] * ratio + 1.0f + 2.0f * (float)y,
                               (float)id[2] * ratio + 1.0f + 2.0f * (float)z);
        sum.x += v.x;
        sum.y += v.y;
        sum.z += v.z;
        sum.w += v.w;
      }
  if (blocks == 1) return sum;
  float n = (float)(blocks * blocks * blocks);
  return make_float4(sum.x / n, sum.y / n, sum.z / n, sum.w / n);
}

// smokeVelocityAverage (smoke.wgsl:4-21): the coarse smoke under velocity cell `p`, for
// buoyancy. With velocity cells at most as wide as smoke cells, one sample at the cell's
// center; with velocity cells two smoke cells wide, their box average.
__device__ __forceinline__ float smokeVelocityAverage(const Ctx &c, const StepArgs &s, const int p[3]) {
  const Params &P = s.P;
  float sp[3];
  if (P.smokeRatio == 1) {
    for (int a = 0; a < 3; a++) sp[a] = (((float)p[a] + 0.5f) * P.hv) * P.invHs;
  } else {
    for (int a = 0; a < 3; a++) sp[a] = (float)p[a] * 2.0f + 1.0f;
  }
  float lo, hi;
  return sampleSmoke(c, s.p.S, sp[0], sp[1], sp[2], lo, hi);
}

// centerForce (fluid-forces.wgsl:79-104), zero where the brick does not compute.
__device__ __forceinline__ void cellForce(const Ctx &c, const StepArgs &s, const int p[3], float out[3]) {
  const Params &P = s.P;
  const int sV = c.L.sV;
  if (entryAt(c.home, c.dir, p[0] >> sV, p[1] >> sV, p[2] >> sV) < 0) {
    out[0] = out[1] = out[2] = 0.f;
    return;
  }
  float w[3] = {((float)p[0] + 0.5f) * P.hv, ((float)p[1] + 0.5f) * P.hv, ((float)p[2] + 0.5f) * P.hv};
  if (solidAtNear(c, P, s.p.solids, w[0], w[1], w[2])) {
    out[0] = out[1] = out[2] = 0.f;
    return;
  }
  int d = c.L.B / c.L.b;
  float4 f = velocityCellAverage(c, s.p.F, d, p);
  float conf[3] = {0.f, 0.f, 0.f};
  if (P.vorticity != 0.f) {
    float4 cc = curlCellNear(c, s.p.Curl, p[0], p[1], p[2]);
    float eta[3] = {curlCellNear(c, s.p.Curl, p[0] + 1, p[1], p[2]).w - curlCellNear(c, s.p.Curl, p[0] - 1, p[1], p[2]).w,
                    curlCellNear(c, s.p.Curl, p[0], p[1] + 1, p[2]).w - curlCellNear(c, s.p.Curl, p[0], p[1] - 1, p[2]).w,
                    curlCellNear(c, s.p.Curl, p[0], p[1], p[2] + 1).w - curlCellNear(c, s.p.Curl, p[0], p[1], p[2] - 1).w};
    float len = length3(eta[0], eta[1], eta[2]) + 0.0001f;
    float n[3] = {eta[0] / len, eta[1] / len, eta[2] / len};
    float cr[3] = {n[1] * cc.z - n[2] * cc.y, n[2] * cc.x - n[0] * cc.z, n[0] * cc.y - n[1] * cc.x};
    for (int a = 0; a < 3; a++) conf[a] = cr[a] * P.vorticity * P.hv;
  }
  float ext[3];
  externalForces(s, p, w, ext);
  out[0] = conf[0] + ext[0] + 0.f;
  const float density = P.coarse ? smokeVelocityAverage(c, s, p) : f.x;
  out[1] = conf[1] + ext[1] + (P.buoyancy * f.y - P.smokeWeight * density);
  out[2] = conf[2] + ext[2] + 0.f;
}

// sourceVelocity (fluid-forces.wgsl:122-149).
__device__ __forceinline__ float sourceVelocity(const StepArgs &s, float value, const float face[3], int comp) {
  float result = value;
  for (int i = 0; i < s.P.emitterCount; i++) {
    const Emitter &e = s.emitters[i];
    if (e.response == 0.f) continue;
    float wgt = emitterWeight(e, face);
    if (wgt == 0.f) continue;
    float radial = 0.f;
    if (e.outward != 0.f) {
      float d0 = face[0] - e.pos[0], d1 = face[1] - e.pos[1], d2 = face[2] - e.pos[2];
      float dc = comp == 0 ? d0 : comp == 1 ? d1 : d2;
      radial = dc / fmaxv(length3(d0, d1, d2), 1e-6f) * e.outward;
    }
    float vc = comp == 0 ? e.vel[0] : comp == 1 ? e.vel[1] : e.vel[2];
    result = mixv(result, vc + radial, 1.0f - fexp(-e.response * wgt * s.P.dt));
  }
  if (s.P.meshRecords) {
    const float4 motion = s.meshRecords[2 * meshRecordAt(s, face) + 1];
    if (motion.w > 0.f) {
      const float mc = comp == 0 ? motion.x : comp == 1 ? motion.y : motion.z;
      result = mixv(result, mc / motion.w, 1.0f - fexp(-motion.w * s.P.dt));
    }
  }
  return result;
}

// The force at each velocity cell of the active bricks (centerForce, fluid-forces.wgsl:
// 79-118), evaluated once per cell into `out` ([slot][axis][cell]); applyForces reads a
// cell's and its +x/+y/+z neighbors' (zero where a brick does not compute).
__device__ __forceinline__ void cell_forces_body(const StepArgs &s, const int vb) {
  float *out = s.p.Vadv;
  BRICK_PREAMBLE(1)
  const int nV = s.L.nV;
  float *o = out + slotBase(slot * 3, s.L.sV);
  for (int l = threadIdx.x; l < nV; l += blockDim.x) {
    int p[3];
    velCellCoords(c, l, p);
    float f[3];
    cellForce(c, s, p, f);
    o[l] = f[0];
    o[nV + l] = f[1];
    o[2 * nV + l] = f[2];
  }
}
FP_KERNEL_PAIR(cell_forces, 256, StepArgs)


// applyForces (fluid-forces.wgsl:151-182): cell forces averaged onto the faces, damping,
// emitter velocity coupling and closed solid faces.
__device__ __forceinline__ void forces_body(const StepArgs &s, const int vb) {
  const float *cellF = s.p.Vadv;
  BRICK_PREAMBLE(1)
  const Params &P = s.P;
  const int nV = s.L.nV;
  const float damp = P.damp;
  const float *vc = s.p.Vcor + slotBase(slot * 3, s.L.sV);
  const float *fh = cellF + slotBase(slot * 3, s.L.sV);
  float *o = s.p.Vf + slotBase(slot * 3, s.L.sV);
  for (int l = threadIdx.x; l < nV; l += blockDim.x) {
    int p[3];
    velCellCoords(c, l, p);
    float force[3] = {0.5f * (fh[l] + velCellNear(c, cellF, 0, p[0] + 1, p[1], p[2])),
                      0.5f * (fh[nV + l] + velCellNear(c, cellF, 1, p[0], p[1] + 1, p[2])),
                      0.5f * (fh[2 * nV + l] + velCellNear(c, cellF, 2, p[0], p[1], p[2] + 1))};
    float pos[3] = {((float)p[0] + 0.5f) * P.hv, ((float)p[1] + 0.5f) * P.hv, ((float)p[2] + 0.5f) * P.hv};
    float v[3];
    for (int a = 0; a < 3; a++) v[a] = vc[a * nV + l] * damp + P.dt * force[a];
    float half = 0.5f * P.hv;
    float fx[3] = {pos[0] + half, pos[1], pos[2]}, fy[3] = {pos[0], pos[1] + half, pos[2]},
          fz[3] = {pos[0], pos[1], pos[2] + half};
    float out0 = sourceVelocity(s, v[0], fx, 0), out1 = sourceVelocity(s, v[1], fy, 1),
          out2 = sourceVelocity(s, v[2], fz, 2);
    if (solidAtNear(c, P, s.p.solids, pos[0], pos[1], pos[2])) {
      o[l] = o[nV + l] = o[2 * nV + l] = 0.f;
      continue;
    }
    if (solidAtNear(c, P, s.p.solids, pos[0] + P.hv, pos[1], pos[2])) out0 = 0.f;
    if (solidAtNear(c, P, s.p.solids, pos[0], pos[1] + P.hv, pos[2])) out1 = 0.f;
    if (solidAtNear(c, P, s.p.solids, pos[0], pos[1], pos[2] + P.hv)) out2 = 0.f;
    o[l] = out0;
    o[nV + l] = out1;
    o[2 * nV + l] = out2;
  }
}
FP_KERNEL_PAIR(forces, 256, StepArgs)


// computeDivergence (fluid-forces.wgsl:184-203): rhs = h^2 (div u - expansion).
__device__ __forceinline__ void divergence_body(const StepArgs &s, const int vb) {
  float *rhs0 = s.p.D;
  BRICK_PREAMBLE(1)
  const Params &P = s.P;
  const int nV = s.L.nV;
  const float *vf = s.p.Vf + slotBase(slot * 3, s.L.sV);
  const int d = s.L.B / s.L.b;
  for (int l = threadIdx.x; l < nV; l += blockDim.x) {
    int p[3];
    velCellCoords(c, l, p);
    float h = P.hv;
    float div = (vf[l] - velCellNear(c, s.p.Vf, 0, p[0] - 1, p[1], p[2])) * P.invHv +
                (vf[nV + l] - velCellNear(c, s.p.Vf, 1, p[0], p[1] - 1, p[2])) * P.invHv +
                (vf[2 * nV + l] - velCellNear(c, s.p.Vf, 2, p[0], p[1], p[2] - 1)) * P.invHv;
    float avg;
    if (d == 1) {
      avg = fieldScalarNear(c, s.p.E, p[0], p[1], p[2]);
    } else {
      int blocks = d / 2;
      float ratio = (float)d;
      float sum = 0.f;
      for (int z = 0; z < blocks; z++)
        for (int y = 0; y < blocks; y++)
          for (int x = 0; x < blocks; x++) {
            float cx = (float)p[0] * ratio + 1.0f + 2.0f * (float)x, cy = (float)p[1] * ratio + 1.0f + 2.0f * (float)y,
                  cz = (float)p[2] * ratio + 1.0f + 2.0f * (float)z;
            float qx = cx - 0.5f, qy = cy - 0.5f, qz = cz - 0.5f;
            float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
            int bx = (int)flx, by = (int)fly, bz = (int)flz;
            float fx = qx - flx, fy = qy - fly, fz = qz - flz;
            float v = 0.f;
            for (int zz = 0; zz < 2; zz++)
              for (int yy = 0; yy < 2; yy++)
                for (int xx = 0; xx < 2; xx++) {
                  float w = (xx ? fx : 1.0f - fx) * (yy ? fy : 1.0f - fy) * (zz ? fz : 1.0f - fz);
                  v += fieldScalarNear(c, s.p.E, bx + xx, by + yy, bz + zz) * w;
                }
            sum += v;
          }
      avg = blocks == 1 ? sum : sum / (float)(blocks * blocks * blocks);
    }
    rhs0[slotBase(slot, s.L.sV) + l] = h * h * (div - avg);
  }
}
FP_KERNEL_PAIR(divergence, 256, StepArgs)


// project (fluid-forces.wgsl:205-262).
__device__ __forceinline__ bool pressureUnknown(const Ctx &c, const Params &P, const int p[3]) {
  const int sV = c.L.sV;
  int lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = (p[a] - 1) >> sV;
    hi[a] = p[a] >> sV;
  }
  if (P.ground) lo[1] = lo[1] > 0 ? lo[1] : 0;
  // The loops' early exits pay here (bricksPresent made k_project 27 -> 39 us).
  for (int z = lo[2]; z <= hi[2]; z++)
    for (int y = lo[1]; y <= hi[1]; y++)
      for (int x = lo[0]; x <= hi[0]; x++)
        if (entryAt(c.home, c.dir, x, y, z) < 0) return false;
  return true;
}

__device__ __forceinline__ void project_body(const StepArgs &s, const int vb) {
  const float *pressure = s.pressure;
  BRICK_PREAMBLE(1)
  const Params &P = s.P;
  const int nV = s.L.nV;
  const float *vf = s.p.Vf + slotBase(slot * 3, s.L.sV);
  float *o = s.p.V + slotBase(slot * 3, s.L.sV);
  for (int l = threadIdx.x; l < nV; l += blockDim.x) {
    int p[3];
    velCellCoords(c, l, p);
    float pos[3] = {((float)p[0] + 0.5f) * P.hv, ((float)p[1] + 0.5f) * P.hv, ((float)p[2] + 0.5f) * P.hv};
    if (solidAtNear(c, P, s.p.solids, pos[0], pos[1], pos[2])) {
      o[l] = o[nV + l] = o[2 * nV + l] = 0.f;
      continue;
    }
    float cp = pressure[slotBase(slot, s.L.sV) + l];
    const float cpLo = s.pressureLo ? s.pressureLo[slotBase(slot, s.L.sV) + l] : 0.f;
    bool here = pressureUnknown(c, P, p);
    for (int a = 0; a < 3; a++) {
      int nb[3] = {p[0], p[1], p[2]};
      nb[a] += 1;
      float npos[3] = {((float)nb[0] + 0.5f) * P.hv, ((float)nb[1] + 0.5f) * P.hv, ((float)nb[2] + 0.5f) * P.hv};
      float value = vf[a * nV + l];
      float res;
      if (solidAtNear(c, P, s.p.solids, npos[0], npos[1], npos[2])) res = 0.f;
      else if (!here && !pressureUnknown(c, P, nb)) res = 0.f;
      else if (!s.pressureLo) res = value - (velScalarNear(c, pressure, nb[0], nb[1], nb[2]) - cp) * P.invHv;
      else {
        // A pair hi + lo: the high parts' difference (exact for neighbors within a factor
        // of two) plus the low parts'.
        const float g = (velScalarNear(c, pressure, nb[0], nb[1], nb[2]) - cp) +
                        (velScalarNear(c, s.pressureLo, nb[0], nb[1], nb[2]) - cpLo);
        res = value - g * P.invHv;
      }
      o[a * nV + l] = res;
    }
  }
}
FP_KERNEL_PAIR(project, 256, StepArgs)




// ---- coarse smoke (smoke.wgsl) ----------------------------------------------------------
// With smokeDivisor 2 or 4 smoke density lives on its own grid, sd field cells wide, in
// pools of (B / sd)^3 cells per slot. evolveFields stages each step's smoke production in
// the fields' X channel; these kernels transport the density with the step's velocity, add
// the production's box average and decay it. One 64-lane workgroup covers one brick.

__device__ __forceinline__ void smokeCellCoords(const Ctx &c, int l, int out[3]) {
  const int n = c.L.bs, sh = c.L.sS;
  out[0] = c.home.bx * n + (l & (n - 1));
  out[1] = c.home.by * n + ((l >> sh) & (n - 1));
  out[2] = c.home.bz * n + (l >> (2 * sh));
}

// advectSmoke (smoke.wgsl:104-126): the backward trace from each smoke cell's center and
// the predictor. For the corrector it leaves the cell's velocity and the range of the
// predictor's eight donors in the scalar backtrace planes, which evolveFields has finished
// with: (velocity, donor minimum) then (donor maximum, 0, 0, 0) per smoke cell.
__device__ __forceinline__ void advect_smoke_body(const StepArgs &s, const int vb) {
  BRICK_PREAMBLE(1)
  const Params &P = s.P;
  const int nS = s.L.nS;
  float4 *don = s.p.Don + slotBase(slot * 3, s.L.sF);
  for (int l = threadIdx.x; l < nS; l += blockDim.x) {
    int cc[3];
    smokeCellCoords(c, l, cc);
    const float pos[3] = {((float)cc[0] + 0.5f) * P.hs, ((float)cc[1] + 0.5f) * P.hs, ((float)cc[2] + 0.5f) * P.hs};
    float density = 0.f, lo = 0.f, hi = 0.f;
    float cv[3] = {0.f, 0.f, 0.f};
    if (!solidAtNear(c, P, s.p.solids, pos[0], pos[1], pos[2])) {
      float dep[3];
      velocityAt(c, P, s.p.V, pos[0], pos[1], pos[2], cv);
      traceBack(c, P, s.p.V, pos, cv, dep);
      density = sampleSmoke(c, s.p.S, dep[0] * P.invHs, dep[1] * P.invHs, dep[2] * P.invHs, lo, hi);
    }
    s.p.Sp[slotBase(slot, s.L.sS) + l] = density;
    don[l] = make_float4(cv[0], cv[1], cv[2], lo);
    don[nS + l] = make_float4(hi, 0.f, 0.f, 0.f);
  }
}
FP_KERNEL_PAIR(advect_smoke, 64, StepArgs)

// evolveSmoke (smoke.wgsl:128-171): the MacCormack correction bounded by the donors, decay,
// the box average of the fine production (already decayed over half a step), the dormant
// cleanup, and content with its support on the smoke grid.
__device__ __forceinline__ void evolve_smoke_body(const StepArgs &s, const int vb) {
  BRICK_PREAMBLE(1)
  __shared__ uint32_t busy;
  if (threadIdx.x == 0) busy = 0u;
  __syncthreads();
  const Params &P = s.P;
  const int nS = s.L.nS;
  const float4 *don = s.p.Don + slotBase(slot * 3, s.L.sF);
  const float ratio = P.hv / P.hs;
  uint32_t myBusy = 0u;
  for (int l = threadIdx.x; l < nS; l += blockDim.x) {
    int cc[3];
    smokeCellCoords(c, l, cc);
    const float pos[3] = {((float)cc[0] + 0.5f) * P.hs, ((float)cc[1] + 0.5f) * P.hs, ((float)cc[2] + 0.5f) * P.hs};
    const size_t at = slotBase(slot, s.L.sS) + l;
    float density = 0.f;
    if (!solidAtNear(c, P, s.p.solids, pos[0], pos[1], pos[2])) {
      const float4 d0 = don[l];
      const float cv[3] = {d0.x, d0.y, d0.z};
      const float predictor = s.p.Sp[at];
      float bounded = predictor;
      if (P.maccormack) {
        const float lo = d0.w, hi = don[nS + l].x;
        if (lo != hi || predictor != lo) {
          float fw[3], rlo, rhi;
          traceForward(c, P, s.p.V, pos, cv, fw);
          const float reversed = sampleSmoke(c, s.p.Sp, fw[0] * P.invHs, fw[1] * P.invHs, fw[2] * P.invHs, rlo, rhi);
          const float corrected = predictor + 0.5f * (s.p.S[at] - reversed);
          bounded = (corrected < lo || corrected > hi) ? predictor : corrected;
        }
      }
      const float4 produced = velocityCellAverage(c, s.p.F, s.L.sd, cc);
      density = fmaxv(0.f, bounded * P.lossS + produced.x);
      if (produced.x == 0.f && produced.z == 0.f && fmaxv(density, fmaxv(produced.y, produced.w)) < 0.001f)
        density = 0.f;
      if (density > P.cutoff) myBusy |= supportOn(s.L.bs, ratio, P.invHs, P.dt, cc, cv);
    }
    s.p.Sn[at] = density;
  }
  if (myBusy) atomicOr(&busy, myBusy);
  __syncthreads();
  if (threadIdx.x == 0 && busy) {
    uint32_t local = (uint32_t)id & 63u;
    uint32_t bits = (1u << (local & 3u)) | (16u << ((local >> 2) & 3u)) | (256u << (local >> 4));
    atomicOr(&s.activity[s.brickCount + (id >> 6)], bits);
    atomicOr(&s.activity[id], CONTENT | (busy << 3));
  }
}
FP_KERNEL_PAIR(evolve_smoke, 64, StepArgs)

}  // namespace fp
