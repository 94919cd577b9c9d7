// ORACLE — test infrastructure only (see solver.hpp). A C ABI over the CPU restatement,
// loaded with ctypes by the tests, smoke() and bench.py's CPU baseline. Entry points mirror
// include/firepro.h's, prefixed orc_, with the arithmetic width chosen at creation.
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <type_traits>
#include <string>

#include "render.hpp"
#include "solver.hpp"

namespace {
thread_local std::string g_error;

struct Base {
  virtual ~Base() = default;
  virtual void step(const fp_frame &f, double dt) = 0;
  virtual void reset() = 0;
  virtual void stats(fp_stats *s) = 0;
  virtual void tables(fp_tables *t) = 0;
  virtual int state(int32_t *ids, float *fields, float *vel, float *pressure, float *expansion, int cap) = 0;
  virtual int state64(int32_t *ids, double *fields, double *vel, double *pressure, double *expansion, int cap) = 0;
  virtual void write(const float *fields, const float *vel) = 0;
  virtual int smoke(int32_t *ids, float *density, int cap) = 0;
  virtual void set(double cutoff, double budget, int ground, int which) = 0;
  virtual const std::vector<float> &noise() = 0;
  virtual void contentTiles(int32_t *out, int *count) = 0;
  virtual void writeActivity(const uint32_t *brickFlags, int n, const int32_t *tiles, int tileCount) = 0;
  virtual void readActivity(uint32_t *brickFlags, int n) = 0;
  virtual int tileTables(int32_t *records, int32_t *hash, int cap, int hashCap) = 0;
  virtual int brickCoords(const int32_t *ids, int n, int32_t *out) = 0;
  virtual int rhs(double *out, int cap) = 0;
  virtual void setPressure(int mode, double tol, int maxCycles) = 0;
  virtual void pressureReport(int *cycles, double *residual) = 0;
  virtual int render(const fp_camera &cam, const fp_render_settings &rs, float *rgba) = 0;
  virtual int readRender(fp_render_tables *t) = 0;
  uint64_t renderCounts[4] = {0, 0, 0, 0};  // the latest image's sample counts
  virtual int probe(const fp_camera &cam, const fp_render_settings &rs, int what, const float *in, float *out) = 0;
  virtual int sceneLights(const fp_render_settings &rs, const double *xyz, int count, float *sums) = 0;
  virtual int stage(int which, double *out, int cap) = 0;
  virtual int forceProbe(const int32_t *cells, int n, double *out) = 0;
};

template <class R> struct Impl : Base {
  oracle::Sim<R> sim;
  explicit Impl(const fp_grid &g) : sim(g) {}
  void step(const fp_frame &f, double dt) override { sim.step(f, dt); }
  void reset() override {
    sim.reset();
    cache.invalidate();
  }
  void stats(fp_stats *s) override {
    s->time = sim.time;
    s->steps = sim.steps;
    s->active_bricks = sim.activeBrickCount;
    s->active_voxels = sim.activeBrickCount * (int64_t)sim.nF;
    s->pool_slots = sim.poolSlots;
    s->used_slots = sim.usedSlots;
    s->pool_limited = sim.poolLimited;
    s->tile_capacity = sim.capacity;
    s->tile_count = sim.dir.count();
    s->pool_side = sim.side;
    s->memory_bytes = 0;
  }
  void tables(fp_tables *t) override {
    int n = sim.brickCount();
    t->brick_count = n;
    t->pool_slots = sim.poolSlots;
    t->free_count = sim.freeCount;
    t->failures = (int32_t)sim.lastFailures;
    if (t->pages)
      for (int i = 0; i < n; i++) t->pages[i] = sim.pages[i];
    if (t->entries)
      for (int i = 0; i < n; i++) t->entries[i] = sim.entries[i];
    if (t->records)
      for (int i = 0; i < n; i++)
        for (int k = 0; k < 28; k++)
          t->records[(size_t)i * 28 + k] = sim.entries[i] >= 0 ? sim.records[(size_t)i * 28 + k] : 0;
    if (t->free_slots)
      for (int i = 0; i < sim.poolSlots; i++) t->free_slots[i] = sim.freeSlots[i];
    t->active_count = (int32_t)sim.activeList.size();
    if (t->active)
      for (size_t i = 0; i < sim.activeList.size(); i++) t->active[i] = sim.activeList[i];
    std::vector<std::pair<int32_t, uint32_t>> z = sim.zeroList;
    std::sort(z.begin(), z.end());
    t->zeroed_count = (int32_t)z.size();
    if (t->zeroed)
      for (size_t i = 0; i < z.size(); i++) {
        t->zeroed[2 * i] = z[i].first;
        t->zeroed[2 * i + 1] = (int32_t)z[i].second;
      }
  }
  template <class T> int stateT(int32_t *ids, T *fields, T *vel, T *pressure, T *expansion, int cap) {
    int n = (int)sim.activeList.size();
    if (n > cap) return -1;
    for (int i = 0; i < n; i++) {
      int id = sim.activeList[i];
      if (ids) ids[i] = id;
      int s = sim.pageSlot(sim.entries[id] & 0xffffff);
      if (fields)
        for (int c = 0; c < sim.nF * 4; c++) fields[(size_t)i * sim.nF * 4 + c] = (T)sim.F[(size_t)s * sim.nF * 4 + c];
      if (expansion)
        for (int c = 0; c < sim.nF; c++) expansion[(size_t)i * sim.nF + c] = (T)sim.E[(size_t)s * sim.nF + c];
      if (vel)
        for (int c = 0; c < sim.nV * 3; c++) vel[(size_t)i * sim.nV * 3 + c] = (T)sim.V[(size_t)s * sim.nV * 3 + c];
      if (pressure)
        for (int c = 0; c < sim.nV; c++) pressure[(size_t)i * sim.nV + c] = (T)sim.Pfine[(size_t)s * sim.nV + c];
    }
    return n;
  }
  int state(int32_t *ids, float *f, float *v, float *p, float *e, int cap) override { return stateT(ids, f, v, p, e, cap); }
  int state64(int32_t *ids, double *f, double *v, double *p, double *e, int cap) override {
    return stateT(ids, f, v, p, e, cap);
  }
  // The coarse smoke density of each active brick, in active-list order.
  int smoke(int32_t *ids, float *density, int cap) override {
    if (!sim.coarse()) throw oracle::Error("this simulation has no coarse smoke grid (smokeDivisor 1).");
    int n = (int)sim.activeList.size();
    if (n > cap) return -1;
    for (int i = 0; i < n; i++) {
      int id = sim.activeList[i];
      if (ids) ids[i] = id;
      int s = sim.pageSlot(sim.entries[id] & 0xffffff);
      if (density)
        for (int c = 0; c < sim.nS; c++) density[(size_t)i * sim.nS + c] = (float)sim.S[(size_t)s * sim.nS + c];
    }
    return n;
  }
  void write(const float *fields, const float *vel) override {
    int n = (int)sim.activeList.size();
    for (int i = 0; i < n; i++) {
      int id = sim.activeList[i];
      int s = sim.pageSlot(sim.entries[id] & 0xffffff);
      if (fields)
        for (int c = 0; c < sim.nF * 4; c++) sim.F[(size_t)s * sim.nF * 4 + c] = (R)fields[(size_t)i * sim.nF * 4 + c];
      if (vel)
        for (int c = 0; c < sim.nV * 3; c++) sim.V[(size_t)s * sim.nV * 3 + c] = (R)vel[(size_t)i * sim.nV * 3 + c];
    }
  }
  void set(double cutoff, double budget, int ground, int which) override {
    if (which == 0) sim.cutoff = cutoff;
    if (which == 1) sim.voxelBudget = budget;
    if (which == 2) sim.ground = ground != 0;
  }
  const std::vector<float> &noise() override { return sim.noise; }
  void contentTiles(int32_t *out, int *count) override {
    int n = (int)sim.contentTiles.size();
    if (out)
      for (int i = 0; i < std::min(n, *count); i++) {
        out[4 * i] = sim.contentTiles[i].first[0];
        out[4 * i + 1] = sim.contentTiles[i].first[1];
        out[4 * i + 2] = sim.contentTiles[i].first[2];
        out[4 * i + 3] = (int32_t)sim.contentTiles[i].second;
      }
    *count = n;
  }
  void writeActivity(const uint32_t *flags, int n, const int32_t *tiles, int tileCount) override {
    for (int i = 0; i < n && i < (int)sim.activity.size(); i++) sim.activity[i] = flags[i];
    sim.contentTiles.clear();
    for (int i = 0; i < tileCount; i++)
      sim.contentTiles.push_back({{tiles[4 * i], tiles[4 * i + 1], tiles[4 * i + 2]}, (uint32_t)tiles[4 * i + 3]});
  }
  int tileTables(int32_t *records, int32_t *hash, int cap, int hashCap) override {
    const auto &t = sim.tables;
    if ((int)t.records.size() > cap * 32 || (int)t.hash.size() > hashCap * 4) return -1;
    std::copy(t.records.begin(), t.records.end(), records);
    std::copy(t.hash.begin(), t.hash.end(), hash);
    return (int)(t.hash.size() / 4);
  }
  int brickCoords(const int32_t *ids, int n, int32_t *out) override {
    for (int i = 0; i < n; i++) sim.brickOf(ids[i], out + 3 * i);
    return n;
  }
  // Level 0's right-hand side h^2 (div u - expansion) of the last step, per active brick.
  int rhs(double *out, int cap) override {
    int n = (int)sim.activeList.size();
    if (n > cap) return -1;
    for (int i = 0; i < n; i++) {
      int s = sim.pageSlot(sim.entries[sim.activeList[i]] & 0xffffff);
      for (int c = 0; c < sim.nV; c++) out[(size_t)i * sim.nV + c] = (double)sim.D[(size_t)s * sim.nV + c];
    }
    return n;
  }
  // An intermediate array of the last step, per active brick in active-list order, in
  // FP64: 0 the advected velocity, 1 the corrected velocity, 2 the forced velocity (three
  // planes of nV faces each), 3 the curl (four channels per velocity cell), 4 the scalar
  // predictor (four channels per field cell), 5 the MacCormack limiter decisions (one mask
  // per field cell). For pinning tests of the stages and the precision report.
  int stage(int which, double *out, int cap) override {
    const std::vector<R> *buf[] = {&sim.Vadv, &sim.Vcor, &sim.Vf, &sim.Curl, &sim.Fp};
    const size_t per[] = {(size_t)sim.nV * 3, (size_t)sim.nV * 3, (size_t)sim.nV * 3, (size_t)sim.nV * 4,
                          (size_t)sim.nF * 4, (size_t)sim.nF};
    if (which < 0 || which > 5) throw oracle::Error("stage must be 0 to 5.");
    int n = (int)sim.activeList.size();
    if (n > cap) return -1;
    for (int i = 0; i < n; i++) {
      size_t s = (size_t)sim.pageSlot(sim.entries[sim.activeList[i]] & 0xffffff);
      for (size_t c = 0; c < per[which]; c++)
        out[i * per[which] + c] =
            which == 5 ? (double)sim.limiter[s * per[which] + c] : (double)(*buf[which])[s * per[which] + c];
    }
    return n;
  }
  // centerForce (fluid-forces.wgsl:79-104) at velocity cells (x, y, z triples) on the last
  // step's fields and curl: confinement, external forces and buoyancy, in FP64.
  int forceProbe(const int32_t *cells, int n, double *out) override {
    for (int i = 0; i < n; i++) {
      R f[3];
      sim.centerForce(cells + 3 * i, f);
      for (int a = 0; a < 3; a++) out[3 * i + a] = (double)f[a];
    }
    return n;
  }
  void readActivity(uint32_t *flags, int n) override {
    for (int i = 0; i < n && i < (int)sim.activity.size(); i++) flags[i] = sim.activity[i];
  }

  void setPressure(int mode, double tol, int maxCycles) override {
    if (mode < 0 || mode > 3) throw oracle::Error("pressure mode must be 0 to 3.");
    sim.pressureMode = mode;
    sim.pressureTol = tol;
    sim.pressureMaxCycles = maxCycles;
  }
  void pressureReport(int *cycles, double *residual) override {
    if (cycles) *cycles = sim.lastCycles;
    if (residual) *residual = sim.lastResidual;
  }

  // ---- render (single precision only) ----
  oracle::RenderCache cache;
  int lastWork = -1;
  void checkRender(const fp_render_settings &rs) {
    if constexpr (!std::is_same<R, float>::value) throw std::runtime_error("the render oracle is single precision");
    int ld = rs.lighting_divisor;
    if (ld != 1 && ld != 2 && ld != 4) throw oracle::Error("lightingDivisor must be 1, 2 or 4.");
    if (sim.B / ld < 1) throw oracle::Error("lightingDivisor exceeds the brick size.");
  }
  int render(const fp_camera &cam, const fp_render_settings &rs, float *rgba) override {
    checkRender(rs);
    if constexpr (std::is_same<R, float>::value) {
      auto p = oracle::renderParams(sim.voxel, sim.B, cam, rs, sim.time);
      lastWork = cache.prepare(sim, p);
      if (rgba) cache.view(sim, p).image(rgba, this->renderCounts);
    }
    return 0;
  }
  int readRender(fp_render_tables *t) override {
    if constexpr (std::is_same<R, float>::value) {
      t->light_work = lastWork;
      t->box_count = (int32_t)cache.boxes.size();
      if (t->boxes)
        for (int i = 0; i < std::min(t->box_count, t->box_capacity); i++) {
          const auto &b = cache.boxes[i];
          int32_t w[8] = {b.origin[0], b.origin[1], b.origin[2], b.first, b.size[0], b.size[1], b.size[2], 0};
          std::copy(w, w + 8, t->boxes + 8 * i);
        }
      int pages = 0;
      for (const auto &b : cache.boxes) pages += b.size[0] * b.size[1] * b.size[2];
      t->page_count = pages;
      if (t->page_table)
        for (int i = 0; i < std::min(pages, t->page_capacity); i++) t->page_table[i] = cache.table[i];
      int E = sim.B + 1, LE = (1 << (sim.sF - (cache.lightDivisor == 4 ? 2 : cache.lightDivisor == 2 ? 1 : 0))) + 1;
      int nO = (sim.B / 8) * (sim.B / 8) * (sim.B / 8);
      size_t nR = (size_t)E * E * E * 4, nL = (size_t)LE * LE * LE * 4;
      oracle::RenderParams p{};
      p.lightShift = sim.sF;
      p.lightDivisor = 1;
      auto view = cache.view(sim, p);
      int n = 0;
      for (int id : sim.activeList) {
        int page, bi, brick[3];
        if (!view.slotted(id, page, bi, brick)) continue;
        if (n < t->brick_capacity) {
          int s = sim.pageSlot(page);
          if (t->bricks) t->bricks[n] = id;
          if (t->field) std::copy(&cache.R[s * nR], &cache.R[s * nR] + nR, t->field + n * nR);
          if (t->light) std::copy(&cache.L[s * nL], &cache.L[s * nL] + nL, t->light + n * nL);
          if (t->occupancy) std::copy(&cache.O[(size_t)s * nO], &cache.O[(size_t)s * nO] + nO, t->occupancy + (size_t)n * nO);
          if (t->smoke && sim.coarse()) {
            size_t nRS = (size_t)(sim.bs + 1) * (sim.bs + 1) * (sim.bs + 1);
            std::copy(&cache.RS[s * nRS], &cache.RS[s * nRS] + nRS, t->smoke + n * nRS);
          }
        }
        n++;
      }
      t->brick_count = n;
      if (t->blackbody) std::copy(cache.lut.begin(), cache.lut.end(), t->blackbody);
      return 0;
    }
    return -1;
  }
  int sceneLights(const fp_render_settings &rs, const double *xyz, int count, float *sums) override {
    checkRender(rs);
    if constexpr (std::is_same<R, float>::value) {
      fp_camera cam{};
      cam.width = cam.height = 1;
      auto p = oracle::renderParams(sim.voxel, sim.B, cam, rs, sim.time);
      std::vector<std::array<float, 3>> anchors(count);
      for (int i = 0; i < count; i++)
        for (int a = 0; a < 3; a++) anchors[i][a] = (float)xyz[3 * i + a];
      if (sim.steps == 0) {
        std::fill(sums, sums + (size_t)8 * count, 0.0f);
        return 0;
      }
      oracle::sceneLights(sim, cache.view(sim, p), anchors, sums);
    }
    return 0;
  }
  // Single functions of the render restatement on the latest prepared state, for tests:
  // 0 reconstruct(cell xyz, mode) in box 0; 1 incidentLight(world xyz) in box 0;
  // 2 the jitter of pixel (x, y) at time t; 3 fexp2(x).
  int probe(const fp_camera &cam, const fp_render_settings &rs, int what, const float *in, float *out) override {
    if constexpr (std::is_same<R, float>::value) {
      auto p = oracle::renderParams(sim.voxel, sim.B, cam, rs, sim.time);
      auto view = cache.view(sim, p);
      if (what == 2) {
        out[0] = oracle::Renderer::jitter((uint32_t)in[0], (uint32_t)in[1], in[2]);
        return 0;
      }
      if (what == 3) {
        out[0] = oracle::fexp2(in[0]);
        return 0;
      }
      if (cache.boxes.empty()) return -1;
      oracle::Vec4 v{};
      if (what == 0) v = view.reconstruct(cache.boxes[0], in, in[3]);
      else if (what == 1) v = view.incidentLight(cache.boxes[0], in);
      else return -1;
      out[0] = v.x;
      out[1] = v.y;
      out[2] = v.z;
      out[3] = v.w;
      return 0;
    }
    return -1;
  }
};

template <class F> int guard(F fn) {
  try {
    fn();
    return FP_OK;
  } catch (const oracle::Error &e) {
    g_error = e.what();
    return FP_EINVAL;
  } catch (const std::exception &e) {
    g_error = e.what();
    return FP_ESTATE;
  }
}
}  // namespace

extern "C" {
const char *orc_last_error(void) { return g_error.c_str(); }

// Host threads for the per-brick passes and the image rows (OpenMP); results do not depend
// on the count. Returns the count in effect.
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int orc_create(const fp_grid *grid, int bits, void **out) {
  return guard([&] {
    if (bits == 64) *out = new Impl<double>(*grid);
    else *out = new Impl<float>(*grid);
  });
}
void orc_destroy(void *h) { delete (Base *)h; }
int orc_step(void *h, const fp_frame *f, double dt) { return guard([&] { ((Base *)h)->step(*f, dt); }); }
int orc_reset(void *h) { return guard([&] { ((Base *)h)->reset(); }); }
int orc_stats(void *h, fp_stats *s) { return guard([&] { ((Base *)h)->stats(s); }); }
int orc_tables(void *h, fp_tables *t) { return guard([&] { ((Base *)h)->tables(t); }); }
int orc_state(void *h, int32_t *ids, float *f, float *v, float *p, float *e, int cap) {
  return ((Base *)h)->state(ids, f, v, p, e, cap);
}
int orc_state64(void *h, int32_t *ids, double *f, double *v, double *p, double *e, int cap) {
  return ((Base *)h)->state64(ids, f, v, p, e, cap);
}
int orc_smoke(void *h, int32_t *ids, float *density, int cap) {
  int n = -1;
  int rc = guard([&] { n = ((Base *)h)->smoke(ids, density, cap); });
  return rc == FP_OK ? n : rc;
}
int orc_write_state(void *h, const float *f, const float *v) { return guard([&] { ((Base *)h)->write(f, v); }); }
int orc_set_cutoff(void *h, double c) { return guard([&] { ((Base *)h)->set(c, 0, 0, 0); }); }
int orc_set_budget(void *h, double v) { return guard([&] { ((Base *)h)->set(0, v, 0, 1); }); }
int orc_set_ground(void *h, int g) { return guard([&] { ((Base *)h)->set(0, 0, g, 2); }); }
int orc_noise(void *h, float *out) {
  const auto &n = ((Base *)h)->noise();
  std::memcpy(out, n.data(), n.size() * sizeof(float));
  return FP_OK;
}
int orc_content_tiles(void *h, int32_t *out, int *count) {
  ((Base *)h)->contentTiles(out, count);
  return FP_OK;
}
int orc_write_activity(void *h, const uint32_t *flags, int n, const int32_t *tiles, int tileCount) {
  return guard([&] { ((Base *)h)->writeActivity(flags, n, tiles, tileCount); });
}
int orc_read_activity(void *h, uint32_t *flags, int n) {
  ((Base *)h)->readActivity(flags, n);
  return FP_OK;
}
float orc_fexp(float x) { return oracle::fexp(x); }
// The multigrid storage rounding: 2 bfloat16, 3 binary16, else none.
float orc_round16(int mode, float x) {
  return mode == 2 ? oracle::roundBF16(x) : mode == 3 ? oracle::roundF16(x) : x;
}
int orc_tile_tables(void *h, int32_t *records, int32_t *hash, int cap, int hashCap) {
  return ((Base *)h)->tileTables(records, hash, cap, hashCap);
}
int orc_brick_coords(void *h, const int32_t *ids, int n, int32_t *out) { return ((Base *)h)->brickCoords(ids, n, out); }
int orc_stage(void *h, int which, double *out, int cap) {
  int r = 0;
  int rc = guard([&] { r = ((Base *)h)->stage(which, out, cap); });
  return rc != FP_OK ? rc : r;
}
int orc_force_probe(void *h, const int32_t *cells, int n, double *out) { return ((Base *)h)->forceProbe(cells, n, out); }
int orc_rhs(void *h, double *out, int cap) { return ((Base *)h)->rhs(out, cap); }
int orc_set_pressure(void *h, int mode, double tol, int max_cycles) {
  return guard([&] { ((Base *)h)->setPressure(mode, tol, max_cycles); });
}
int orc_pressure_report(void *h, int *cycles, double *residual) {
  ((Base *)h)->pressureReport(cycles, residual);
  return FP_OK;
}
int orc_render(void *h, const fp_camera *cam, const fp_render_settings *rs, float *rgba) {
  return guard([&] { ((Base *)h)->render(*cam, *rs, rgba); });
}
int orc_read_render(void *h, fp_render_tables *t) { return ((Base *)h)->readRender(t); }
// The latest image's sample counts: the slices' positions visited and samples shaded, then
// the march's.
int orc_render_counts(void *h, uint64_t *out) {
  for (int k = 0; k < 4; k++) out[k] = ((Base *)h)->renderCounts[k];
  return FP_OK;
}
int orc_scene_lights(void *h, const fp_render_settings *rs, const double *xyz, int count, float *sums) {
  return guard([&] { ((Base *)h)->sceneLights(*rs, xyz, count, sums); });
}
int orc_render_probe(void *h, const fp_camera *cam, const fp_render_settings *rs, int what, const float *in,
                     float *out) {
  int r = -1;
  int s = guard([&] { r = ((Base *)h)->probe(*cam, *rs, what, in, out); });
  return s != FP_OK ? s : r;
}
void orc_blackbody(float *out) {
  auto t = oracle::blackbodyTable();
  std::copy(t.begin(), t.end(), out);
}

// Closed-form hooks: single functions of the restatement, evaluated in double.
static oracle::Params<double> hookParams(const double *knots, const int *counts, double dt, double lifespan,
                                         double smoke, double heat, double expansion) {
  oracle::Params<double> p{};
  for (int ch = 0; ch < 3; ch++) {
    p.knotCount[ch] = counts[ch];
    for (int i = 0; i < counts[ch]; i++) {
      p.knots[ch][i][0] = knots[(ch * 8 + i) * 2];
      p.knots[ch][i][1] = knots[(ch * 8 + i) * 2 + 1];
    }
  }
  p.fdt = dt;
  p.lifespan = lifespan;
  p.fsmoke = smoke;
  p.fheat = heat;
  p.fexpansion = expansion;
  return p;
}
double orc_ramp_integral(const double *knots, const int *counts, int ch, double lo, double hi) {
  auto p = hookParams(knots, counts, 1, 1, 0, 0, 0);
  return oracle::rampIntegral(p, ch, lo, hi);
}
void orc_evaluate_flame(const double *knots, const int *counts, double dt, double lifespan, double smoke, double heat,
                        double expansion, double before, double *out4) {
  auto p = hookParams(knots, counts, dt, lifespan, smoke, heat, expansion);
  double after, prod[3];
  oracle::evaluateFlame(p, before, after, prod);
  out4[0] = after;
  out4[1] = prod[0];
  out4[2] = prod[1];
  out4[3] = prod[2];
}
double orc_emitter_weight(const double *pos, const double *size, int box, const double *w) {
  oracle::Emitter<double> e{};
  for (int a = 0; a < 3; a++) {
    e.pos[a] = pos[a];
    e.size[a] = size[a];
  }
  e.box = box != 0;
  for (int a = 0; a < 3; a++) e.invSize[a] = 1.0 / std::max(e.size[a], 0.03);
  return oracle::emitterWeight(e, w);
}
double orc_burst_perlin(uint32_t seedBits, double x, double y, double z) {
  return oracle::burstPerlin<double>(seedBits, x, y, z);
}
// Time `steps` of an oracle run, in seconds, for the CPU baseline.
double orc_time_steps(void *h, const fp_frame *f, double dt, int steps) {
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < steps; i++) ((Base *)h)->step(*f, dt);
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
}
