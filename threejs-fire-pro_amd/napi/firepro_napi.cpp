// Node N-API binding of the C ABI (include/firepro.h): the engine object Fire Pro's library
// layer constructs (src/library/FireSimulation.ts:203-209) steps through this addon instead
// of encoding WebGPU passes. Plain JavaScript data crosses the boundary: the grid
// (SimulationGrid, src/engine/types.ts:26-36), each step's SimulationFrame
// (types.ts:235-311) and plain results. Failures throw Error with the library's message,
// which keeps the reference's wording (FluidSimulation.ts:362-380, :1368-1369).
//
// Built by __graft_entry__.build() with g++ against Node's own node_api.h; loads
// libfirepro.so from ../firepro through its run path.
#define NAPI_VERSION 4
#include <node_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/firepro.h"

namespace {

// A simulation and the map from the reference's string source ids to the C ABI's
// integer ids (a new id asks for pool slots, as a new string does in the reference).
struct Handle {
  fp_sim *sim = nullptr;
  size_t velocityCells = 0;  // (brickSize / velocityDivisor)^3
  std::unordered_map<std::string, uint64_t> ids;
  uint64_t next = 1;
};

struct JsError {
  std::string message;
};

void check(napi_env env, napi_status st) {
  if (st != napi_ok) {
    const napi_extended_error_info *info = nullptr;
    napi_get_last_error_info(env, &info);
    throw JsError{info && info->error_message ? info->error_message : "N-API call failed"};
  }
}
void fp(int status) {
  if (status != FP_OK) throw JsError{fp_last_error()};
}

// ---- reading JavaScript values ----
bool has(napi_env env, napi_value obj, const char *key) {
  bool out = false;
  check(env, napi_has_named_property(env, obj, key, &out));
  if (!out) return false;
  napi_value v;
  check(env, napi_get_named_property(env, obj, key, &v));
  napi_valuetype t;
  check(env, napi_typeof(env, v, &t));
  return t != napi_undefined && t != napi_null;
}
napi_value get(napi_env env, napi_value obj, const char *key) {
  napi_value v;
  check(env, napi_get_named_property(env, obj, key, &v));
  return v;
}
double number(napi_env env, napi_value v, const char *what) {
  napi_valuetype t;
  check(env, napi_typeof(env, v, &t));
  if (t != napi_number) throw JsError{std::string(what) + " must be a number."};
  double d = 0;
  check(env, napi_get_value_double(env, v, &d));
  return d;
}
double num(napi_env env, napi_value obj, const char *key, double fallback) {
  return has(env, obj, key) ? number(env, get(env, obj, key), key) : fallback;
}
double req(napi_env env, napi_value obj, const char *key) {
  if (!has(env, obj, key)) throw JsError{std::string(key) + " is required."};
  return number(env, get(env, obj, key), key);
}
bool boolean(napi_env env, napi_value obj, const char *key, bool fallback) {
  if (!has(env, obj, key)) return fallback;
  napi_value v = get(env, obj, key);
  napi_valuetype t;
  check(env, napi_typeof(env, v, &t));
  if (t != napi_boolean) throw JsError{std::string(key) + " must be a boolean."};
  bool b = false;
  check(env, napi_get_value_bool(env, v, &b));
  return b;
}
std::string str(napi_env env, napi_value v) {
  size_t n = 0;
  check(env, napi_get_value_string_utf8(env, v, nullptr, 0, &n));
  std::string s(n, '\0');
  check(env, napi_get_value_string_utf8(env, v, &s[0], n + 1, &n));
  return s;
}
uint32_t length(napi_env env, napi_value arr) {
  bool is = false;
  check(env, napi_is_array(env, arr, &is));
  if (!is) throw JsError{"expected an array."};
  uint32_t n = 0;
  check(env, napi_get_array_length(env, arr, &n));
  return n;
}
napi_value at(napi_env env, napi_value arr, uint32_t i) {
  napi_value v;
  check(env, napi_get_element(env, arr, i, &v));
  return v;
}
void vec3(napi_env env, napi_value obj, const char *key, double out[3], const double *fallback = nullptr) {
  if (!has(env, obj, key)) {
    if (!fallback) throw JsError{std::string(key) + " is required."};
    for (int a = 0; a < 3; a++) out[a] = fallback[a];
    return;
  }
  napi_value v = get(env, obj, key);
  if (length(env, v) != 3) throw JsError{std::string(key) + " must have three components."};
  for (uint32_t a = 0; a < 3; a++) out[a] = number(env, at(env, v, a), key);
}
Handle *handle(napi_env env, napi_value v) {
  void *p = nullptr;
  check(env, napi_get_value_external(env, v, &p));
  Handle *h = static_cast<Handle *>(p);
  if (!h || !h->sim) throw JsError{"the simulation has been disposed."};
  return h;
}

// ---- writing JavaScript values ----
napi_value dbl(napi_env env, double d) {
  napi_value v;
  check(env, napi_create_double(env, d, &v));
  return v;
}
void set(napi_env env, napi_value obj, const char *key, napi_value v) {
  check(env, napi_set_named_property(env, obj, key, v));
}
napi_value object(napi_env env) {
  napi_value o;
  check(env, napi_create_object(env, &o));
  return o;
}
napi_value flag(napi_env env, bool b) {
  napi_value v;
  check(env, napi_get_boolean(env, b, &v));
  return v;
}
template <class T> napi_value typed(napi_env env, napi_typedarray_type type, const T *data, size_t count) {
  void *raw = nullptr;
  napi_value buf, arr;
  check(env, napi_create_arraybuffer(env, count * sizeof(T), &raw, &buf));
  if (count) std::memcpy(raw, data, count * sizeof(T));
  check(env, napi_create_typedarray(env, type, count, buf, 0, &arr));
  return arr;
}
napi_value array3(napi_env env, const int32_t *v) {
  napi_value arr;
  check(env, napi_create_array_with_length(env, 3, &arr));
  for (uint32_t a = 0; a < 3; a++) check(env, napi_set_element(env, arr, a, dbl(env, v[a])));
  return arr;
}

// ---- the frame ----
// Storage for one frame's C arrays; `frame` points into it.
struct Frame {
  fp_frame frame{};
  std::vector<fp_source> sources;
  std::vector<std::vector<double>> samples;
  std::vector<fp_force> forces;
  std::vector<std::vector<fp_source>> targets;
  std::vector<fp_collider> colliders;
  std::vector<std::vector<uint32_t>> masks;
};

int shapeOf(const std::string &s) {
  if (s == "sphere") return FP_SHAPE_SPHERE;
  if (s == "box") return FP_SHAPE_BOX;
  if (s == "mesh") return FP_SHAPE_MESH;
  throw JsError{"shape must be sphere, box or mesh."};
}

void source(napi_env env, Handle &h, napi_value o, fp_source &s, std::vector<double> &samples) {
  s = fp_source{};
  std::string id = str(env, get(env, o, "id"));
  auto it = h.ids.find(id);
  if (it == h.ids.end()) it = h.ids.emplace(id, h.next++).first;
  s.id = it->second;
  s.shape = shapeOf(str(env, get(env, o, "shape")));
  vec3(env, o, "position", s.position);
  vec3(env, o, "size", s.size);
  s.flame = req(env, o, "flame");
  s.heat_rate = req(env, o, "heatRate");
  s.smoke_rate = req(env, o, "smokeRate");
  s.fuel_rate = num(env, o, "fuelRate", 0);
  vec3(env, o, "velocity", s.velocity);
  s.velocity_response = req(env, o, "velocityResponse");
  s.outward_speed = req(env, o, "outwardSpeed");
  if (has(env, o, "variation")) {
    napi_value v = get(env, o, "variation");
    s.has_variation = 1;
    s.variation_period = req(env, v, "period");
    s.variation_strength = req(env, v, "strength");
  }
  if (s.shape == FP_SHAPE_MESH && has(env, o, "samples")) {
    napi_value arr = get(env, o, "samples");
    uint32_t n = length(env, arr);
    samples.resize((size_t)n * 4);
    for (uint32_t k = 0; k < n; k++) {
      napi_value e = at(env, arr, k);
      vec3(env, e, "position", &samples[(size_t)k * 4]);
      samples[(size_t)k * 4 + 3] = req(env, e, "weight");
    }
    s.sample_count = (int32_t)n;
    s.surface_area = num(env, o, "surfaceArea", 0);
  }
}

void flame(napi_env env, napi_value o, fp_flame &f) {
  f.lifespan = req(env, o, "lifespan");
  f.smoke_rate = req(env, o, "smokeRate");
  f.heat_rate = req(env, o, "heatRate");
  f.expansion_rate = req(env, o, "expansionRate");
  napi_value ramps = get(env, o, "ramps");
  const char *names[3] = {"smoke", "heat", "expansion"};  // FLAME_CHANNELS (flame.ts:5)
  for (int ch = 0; ch < 3; ch++) {
    napi_value r = get(env, ramps, names[ch]);
    uint32_t n = length(env, r);
    if (n > FP_MAX_FLAME_KNOTS) throw JsError{std::string(names[ch]) + " ramp needs 2–8 knots."};
    f.knot_count[ch] = (int32_t)n;
    for (uint32_t k = 0; k < n; k++) {
      napi_value knot = at(env, r, k);
      if (length(env, knot) != 2) throw JsError{std::string(names[ch]) + " knots are [lifetime, response] pairs."};
      f.knots[ch][k][0] = number(env, at(env, knot, 0), "lifetime");
      f.knots[ch][k][1] = number(env, at(env, knot, 1), "response");
    }
  }
}

void frame(napi_env env, Handle &h, napi_value o, Frame &out) {
  fp_frame &f = out.frame;
  napi_value sim = get(env, o, "simulation");
  f.simulation.buoyancy = req(env, sim, "buoyancy");
  f.simulation.smoke_weight = req(env, sim, "smokeWeight");
  f.simulation.velocity_damping = req(env, sim, "velocityDamping");
  f.simulation.cooling = req(env, sim, "cooling");
  f.simulation.dissipation = req(env, sim, "dissipation");
  f.simulation.vorticity = req(env, sim, "vorticity");
  f.simulation.seed = req(env, sim, "seed");
  flame(env, get(env, o, "combustion"), f.combustion);
  if (has(env, o, "fuel")) {
    napi_value fu = get(env, o, "fuel");
    f.has_fuel = 1;
    f.fuel.enabled = boolean(env, fu, "enabled", false) ? 1 : 0;
    f.fuel.ignition_heat = req(env, fu, "ignitionHeat");
    f.fuel.burn_rate = req(env, fu, "burnRate");
  }
  napi_value srcs = get(env, o, "sources");
  uint32_t ns = length(env, srcs);
  out.sources.resize(ns);
  out.samples.resize(ns);
  for (uint32_t i = 0; i < ns; i++) source(env, h, at(env, srcs, i), out.sources[i], out.samples[i]);
  for (uint32_t i = 0; i < ns; i++)
    if (!out.samples[i].empty()) out.sources[i].samples = out.samples[i].data();
  f.source_count = (int32_t)ns;
  f.sources = out.sources.data();
  if (has(env, o, "forces")) {
    napi_value fs = get(env, o, "forces");
    uint32_t nf = length(env, fs);
    out.forces.resize(nf);
    out.targets.resize(nf);
    std::vector<std::vector<std::vector<double>>> targetSamples(nf);
    for (uint32_t i = 0; i < nf; i++) {
      napi_value e = at(env, fs, i);
      fp_force &fc = out.forces[i];
      fc = fp_force{};
      std::string t = str(env, get(env, e, "type"));
      fc.type = t == "wind" ? FP_FORCE_WIND : t == "turbulence" ? FP_FORCE_TURBULENCE
                : t == "vortex" ? FP_FORCE_VORTEX : t == "radial" ? FP_FORCE_RADIAL : -1;
      if (fc.type < 0) throw JsError{"force.type must be wind, turbulence, vortex or radial."};
      vec3(env, e, "vector", fc.vector);
      vec3(env, e, "center", fc.center);
      fc.strength = req(env, e, "strength");
      fc.scale = req(env, e, "scale");
      fc.lift = req(env, e, "lift");
      fc.inward = req(env, e, "inward");
      if (has(env, e, "sources")) {
        napi_value ts = get(env, e, "sources");
        uint32_t n = length(env, ts);
        out.targets[i].resize(n);
        targetSamples[i].resize(n);
        for (uint32_t k = 0; k < n; k++) source(env, h, at(env, ts, k), out.targets[i][k], targetSamples[i][k]);
        fc.targeted = 1;
        fc.source_count = (int32_t)n;
      }
    }
    // Mesh targets carry their samples; keep them with the frame.
    for (uint32_t i = 0; i < nf; i++)
      for (size_t k = 0; k < out.targets[i].size(); k++)
        if (!targetSamples[i][k].empty()) {
          out.samples.push_back(std::move(targetSamples[i][k]));
          out.targets[i][k].samples = out.samples.back().data();
        }
    for (uint32_t i = 0; i < nf; i++) out.forces[i].sources = out.targets[i].data();
    f.force_count = (int32_t)nf;
    f.forces = out.forces.data();
  }
  napi_value cs = get(env, o, "colliders");
  uint32_t nc = length(env, cs);
  out.colliders.resize(nc);
  out.masks.resize(nc);
  for (uint32_t i = 0; i < nc; i++) {
    napi_value e = at(env, cs, i);
    fp_collider &c = out.colliders[i];
    c = fp_collider{};
    c.shape = shapeOf(str(env, get(env, e, "shape")));
    if (c.shape == FP_SHAPE_MESH) {
      napi_value occ = get(env, e, "occupancy");
      napi_typedarray_type type;
      size_t n = 0, offset = 0;
      void *data = nullptr;
      napi_value ab;
      check(env, napi_get_typedarray_info(env, get(env, occ, "mask"), &type, &n, &data, &ab, &offset));
      if (type != napi_uint32_array) throw JsError{"occupancy.mask must be a Uint32Array."};
      double lo[3], size[3];
      vec3(env, occ, "lo", lo);
      vec3(env, occ, "size", size);
      for (int a = 0; a < 3; a++) {
        c.lo[a] = (int32_t)lo[a];
        c.cells[a] = (int32_t)size[a];
      }
      if ((double)n != size[0] * size[1] * size[2])
        throw JsError{"Mesh collider occupancy does not match its size."};
      out.masks[i].assign((const uint32_t *)data, (const uint32_t *)data + n);
      c.occupancy = out.masks[i].data();
    } else {
      vec3(env, e, "position", c.position);
      vec3(env, e, "size", c.size);
      static const double X[3] = {1, 0, 0}, Y[3] = {0, 1, 0}, Z[3] = {0, 0, 1};
      if (has(env, e, "axes")) {
        napi_value ax = get(env, e, "axes");
        if (length(env, ax) != 3) throw JsError{"collider axes must be three vectors."};
        for (uint32_t k = 0; k < 3; k++) {
          napi_value row = at(env, ax, k);
          if (length(env, row) != 3) throw JsError{"collider axes must be three vectors."};
          for (uint32_t a = 0; a < 3; a++) c.axes[k][a] = number(env, at(env, row, a), "axes");
        }
      } else {
        for (int a = 0; a < 3; a++) {
          c.axes[0][a] = X[a];
          c.axes[1][a] = Y[a];
          c.axes[2][a] = Z[a];
        }
      }
    }
  }
  f.collider_count = (int32_t)nc;
  f.colliders = out.colliders.data();
}

// ---- entry points ----
template <class Fn> napi_value entry(napi_env env, napi_callback_info info, size_t want, Fn fn) {
  std::vector<napi_value> argv(want ? want : 1);
  size_t argc = want;
  try {
    check(env, napi_get_cb_info(env, info, &argc, argv.data(), nullptr, nullptr));
    if (argc < want) throw JsError{"missing arguments."};
    return fn(argv.data());
  } catch (const JsError &e) {
    napi_throw_error(env, nullptr, e.message.c_str());
    return nullptr;
  } catch (const std::exception &e) {
    napi_throw_error(env, nullptr, e.what());
    return nullptr;
  }
}

void finalize(napi_env, void *data, void *) {
  Handle *h = static_cast<Handle *>(data);
  fp_sim_destroy(h->sim);  // NULL after dispose
  delete h;
}

// create(grid) -> handle (FluidSimulation.create, FluidSimulation.ts:359-422)
napi_value Create(napi_env env, napi_callback_info info) {
  return entry(env, info, 1, [&](napi_value *a) {
    fp_grid g{};
    g.voxel_size = has(env, a[0], "voxelSize") ? number(env, get(env, a[0], "voxelSize"), "voxelSize") : NAN;
    g.velocity_divisor = (int32_t)num(env, a[0], "velocityDivisor", 1);
    g.brick_size = (int32_t)num(env, a[0], "brickSize", 16);
    g.smoke_divisor = (int32_t)num(env, a[0], "smokeDivisor", 1);
    if (has(env, a[0], "scalarMacCormack")) {
      napi_value v = get(env, a[0], "scalarMacCormack");
      napi_valuetype t;
      check(env, napi_typeof(env, v, &t));
      if (t != napi_boolean) throw JsError{"scalarMacCormack must be a boolean."};
    }
    g.scalar_maccormack = boolean(env, a[0], "scalarMacCormack", true) ? 1 : 0;
    Handle *h = new Handle();
    if (g.velocity_divisor > 0) {
      const size_t b = (size_t)std::max(1, g.brick_size / g.velocity_divisor);
      h->velocityCells = b * b * b;
    }
    int st = fp_sim_create(&g, &h->sim);
    if (st != FP_OK) {
      delete h;
      fp(st);
    }
    napi_value out;
    check(env, napi_create_external(env, h, finalize, nullptr, &out));
    return out;
  });
}

// step(handle, frame, dt) (FluidSimulation.step, FluidSimulation.ts:1367-1373)
napi_value Step(napi_env env, napi_callback_info info) {
  return entry(env, info, 3, [&](napi_value *a) -> napi_value {
    Handle *h = handle(env, a[0]);
    Frame f;
    frame(env, *h, a[1], f);
    fp(fp_sim_step(h->sim, &f.frame, number(env, a[2], "dt")));
    return nullptr;
  });
}

napi_value Reset(napi_env env, napi_callback_info info) {
  return entry(env, info, 1, [&](napi_value *a) -> napi_value {
    fp(fp_sim_reset(handle(env, a[0])->sim));
    return nullptr;
  });
}

// dispose(handle): idempotent, like FluidSimulation.dispose.
napi_value Dispose(napi_env env, napi_callback_info info) {
  return entry(env, info, 1, [&](napi_value *a) -> napi_value {
    void *p = nullptr;
    check(env, napi_get_value_external(env, a[0], &p));
    Handle *h = static_cast<Handle *>(p);
    if (h && h->sim) {
      fp_sim_destroy(h->sim);
      h->sim = nullptr;
    }
    return nullptr;
  });
}

// setCutoff / setVoxelBudget / setGround (FireSimulation.ts:457-459)
napi_value SetCutoff(napi_env env, napi_callback_info info) {
  return entry(env, info, 2, [&](napi_value *a) -> napi_value {
    fp(fp_sim_set_cutoff(handle(env, a[0])->sim, number(env, a[1], "cutoff")));
    return nullptr;
  });
}
napi_value SetVoxelBudget(napi_env env, napi_callback_info info) {
  return entry(env, info, 2, [&](napi_value *a) -> napi_value {
    fp(fp_sim_set_budget(handle(env, a[0])->sim, number(env, a[1], "voxelBudget")));
    return nullptr;
  });
}
napi_value SetGround(napi_env env, napi_callback_info info) {
  return entry(env, info, 2, [&](napi_value *a) -> napi_value {
    bool g = false;
    check(env, napi_get_value_bool(env, a[1], &g));
    fp(fp_sim_set_ground(handle(env, a[0])->sim, g ? 1 : 0));
    return nullptr;
  });
}

// stats(handle): the getters FireSimulation reads (FireSimulation.ts:296-333).
napi_value Stats(napi_env env, napi_callback_info info) {
  return entry(env, info, 1, [&](napi_value *a) {
    fp_stats s{};
    fp(fp_sim_stats(handle(env, a[0])->sim, &s));
    napi_value o = object(env);
    set(env, o, "time", dbl(env, s.time));
    set(env, o, "steps", dbl(env, (double)s.steps));
    set(env, o, "activeBrickCount", dbl(env, (double)s.active_bricks));
    set(env, o, "activeVoxels", dbl(env, (double)s.active_voxels));
    set(env, o, "poolSlots", dbl(env, (double)s.pool_slots));
    set(env, o, "usedSlots", dbl(env, (double)s.used_slots));
    set(env, o, "poolLimited", flag(env, s.pool_limited != 0));
    set(env, o, "memoryBytes", dbl(env, (double)s.memory_bytes));
    set(env, o, "meshMemoryBytes", dbl(env, (double)s.mesh_memory_bytes));
    set(env, o, "forceMemoryBytes", dbl(env, (double)s.force_memory_bytes));
    set(env, o, "tileCount", dbl(env, (double)s.tile_count));
    return o;
  });
}

// readBricks(handle) -> Float64Array of min corners, 3 per brick (readBricks,
// FluidSimulation.ts:1749-1779); the JavaScript side adds the brick size for the max.
napi_value ReadBricks(napi_env env, napi_callback_info info) {
  return entry(env, info, 1, [&](napi_value *a) {
    fp_sim *sim = handle(env, a[0])->sim;
    fp_stats s{};
    fp(fp_sim_stats(sim, &s));
    int32_t cap = s.tile_capacity * 64;
    std::vector<double> mins((size_t)cap * 3);
    fp(fp_sim_read_bricks(sim, mins.data(), &cap));
    return typed(env, napi_float64_array, mins.data(), (size_t)cap * 3);
  });
}

// showSolverField(handle, 'pressure' | 'divergence' | undefined)
napi_value ShowSolverField(napi_env env, napi_callback_info info) {
  return entry(env, info, 2, [&](napi_value *a) -> napi_value {
    napi_valuetype t;
    check(env, napi_typeof(env, a[1], &t));
    int32_t field = FP_SOLVER_FIELD_NONE;
    if (t == napi_string) {
      std::string s = str(env, a[1]);
      if (s == "pressure") field = FP_SOLVER_FIELD_PRESSURE;
      else if (s == "divergence") field = FP_SOLVER_FIELD_DIVERGENCE;
      else throw JsError{"solver field must be pressure, divergence or undefined."};
    } else if (t != napi_undefined && t != napi_null) {
      throw JsError{"solver field must be pressure, divergence or undefined."};
    }
    fp(fp_sim_show_solver_field(handle(env, a[0])->sim, field));
    return nullptr;
  });
}

// readSolverField(handle) -> {ids: Int32Array, values: Float32Array (velocity cells per brick)}
napi_value ReadSolverField(napi_env env, napi_callback_info info) {
  return entry(env, info, 1, [&](napi_value *a) {
    Handle *h = handle(env, a[0]);
    fp_stats s{};
    fp(fp_sim_stats(h->sim, &s));
    const int32_t cap = (int32_t)std::max<int64_t>(1, s.active_bricks);
    std::vector<int32_t> ids(cap);
    std::vector<float> values((size_t)cap * h->velocityCells);
    int32_t got = fp_sim_read_solver_field(h->sim, ids.data(), values.data(), cap);
    if (got < 0) fp(got);
    napi_value o = object(env);
    set(env, o, "ids", typed(env, napi_int32_array, ids.data(), (size_t)got));
    set(env, o, "values", typed(env, napi_float32_array, values.data(), (size_t)got * h->velocityCells));
    return o;
  });
}

// renderBricks(handle) -> {version, boxes: [{lo, hi}], tiles: Int32Array (4 words per slot)}
napi_value RenderBricks(napi_env env, napi_callback_info info) {
  return entry(env, info, 1, [&](napi_value *a) {
    fp_sim *sim = handle(env, a[0])->sim;
    fp_render_bricks rb{};
    fp(fp_sim_render_bricks(sim, &rb));
    std::vector<int32_t> boxes((size_t)std::max(1, rb.box_count) * 6), tiles((size_t)std::max(1, rb.tile_count) * 4);
    rb.boxes = boxes.data();
    rb.tiles = tiles.data();
    rb.box_capacity = std::max(1, rb.box_count);
    rb.tile_capacity = std::max(1, rb.tile_count);
    fp(fp_sim_render_bricks(sim, &rb));
    napi_value o = object(env), list;
    set(env, o, "version", dbl(env, rb.version));
    check(env, napi_create_array_with_length(env, (size_t)rb.box_count, &list));
    for (int i = 0; i < rb.box_count; i++) {
      napi_value b = object(env);
      set(env, b, "lo", array3(env, &boxes[(size_t)i * 6]));
      set(env, b, "hi", array3(env, &boxes[(size_t)i * 6 + 3]));
      check(env, napi_set_element(env, list, (uint32_t)i, b));
    }
    set(env, o, "boxes", list);
    set(env, o, "tiles", typed(env, napi_int32_array, tiles.data(), (size_t)rb.tile_count * 4));
    set(env, o, "activeCount", dbl(env, rb.active_count));
    return o;
  });
}

// setPressure(handle, 'vcycle' | 'fp32' | 'bf16' | 'fp16', tolerance, maxCycles)
napi_value SetPressure(napi_env env, napi_callback_info info) {
  return entry(env, info, 4, [&](napi_value *a) -> napi_value {
    std::string m = str(env, a[1]);
    int32_t mode = m == "vcycle" ? FP_PRESSURE_VCYCLE : m == "fp32" ? FP_PRESSURE_FP32
                   : m == "bf16" ? FP_PRESSURE_BF16 : m == "fp16" ? FP_PRESSURE_FP16 : -1;
    if (mode < 0) throw JsError{"pressure mode must be vcycle, fp32, bf16 or fp16."};
    fp(fp_sim_set_pressure(handle(env, a[0])->sim, mode, number(env, a[2], "tolerance"),
                           (int32_t)number(env, a[3], "maxCycles")));
    return nullptr;
  });
}

// render(handle, camera, settings) -> Float32Array of width x height premultiplied
// (radiance, 1 - T) (SceneVolumeRenderer.prepare + scene-volume.wgsl:6-37).
napi_value Render(napi_env env, napi_callback_info info) {
  return entry(env, info, 3, [&](napi_value *a) {
    fp_sim *sim = handle(env, a[0])->sim;
    fp_camera cam{};
    napi_value m = get(env, a[1], "inverseViewProjection");
    if (length(env, m) != 16) throw JsError{"inverseViewProjection must have 16 entries."};
    for (uint32_t i = 0; i < 16; i++) cam.inverse_view_projection[i] = number(env, at(env, m, i), "matrix");
    static const double origin[3] = {0, 0, 0};
    vec3(env, a[1], "position", cam.position, origin);
    cam.width = (int32_t)req(env, a[1], "width");
    cam.height = (int32_t)req(env, a[1], "height");
    napi_value o = a[2];
    fp_render_settings rs{};
    rs.fire_intensity = req(env, o, "fireIntensity");
    rs.flame_opacity = req(env, o, "flameOpacity");
    rs.smoke_density = req(env, o, "smokeDensity");
    rs.scattering = req(env, o, "scattering");
    rs.shadow_density = req(env, o, "shadowDensity");
    rs.temperature = req(env, o, "temperature");
    rs.soot_glow = req(env, o, "sootGlow");
    vec3(env, o, "smokeColor", rs.smoke_color);
    vec3(env, o, "fireColor", rs.fire_color);
    vec3(env, o, "sunDirection", rs.sun_direction);
    rs.sun_intensity = req(env, o, "sunIntensity");
    vec3(env, o, "sunColor", rs.sun_color);
    vec3(env, o, "ambient", rs.ambient);
    std::string filt = str(env, get(env, o, "filter"));
    rs.filter = filt == "trilinear" ? 0 : filt == "quadratic" ? 1 : filt == "cubic" ? 2 : -1;
    if (rs.filter < 0) throw JsError{"rendering.filter must be trilinear, quadratic, cubic."};
    rs.ray_steps = (int32_t)req(env, o, "raySteps");
    rs.lighting_divisor = (int32_t)req(env, o, "lightingDivisor");
    rs.half_resolution = boolean(env, o, "halfResolution", true) ? 1 : 0;
    if (cam.width < 1 || cam.height < 1 || cam.width > 16384 || cam.height > 16384)
      throw JsError{"The image must be 1 to 16384 pixels a side."};
    std::vector<float> rgba((size_t)cam.width * cam.height * 4);
    fp(fp_sim_render(sim, &cam, &rs, rgba.data()));
    return typed(env, napi_float32_array, rgba.data(), rgba.size());
  });
}

napi_value Version(napi_env env, napi_callback_info info) {
  return entry(env, info, 0, [&](napi_value *) {
    napi_value v;
    check(env, napi_create_string_utf8(env, fp_version(), NAPI_AUTO_LENGTH, &v));
    return v;
  });
}

napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"create", nullptr, Create, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"step", nullptr, Step, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"reset", nullptr, Reset, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"dispose", nullptr, Dispose, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setCutoff", nullptr, SetCutoff, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setVoxelBudget", nullptr, SetVoxelBudget, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setGround", nullptr, SetGround, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"stats", nullptr, Stats, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"readBricks", nullptr, ReadBricks, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"showSolverField", nullptr, ShowSolverField, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"readSolverField", nullptr, ReadSolverField, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"renderBricks", nullptr, RenderBricks, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"setPressure", nullptr, SetPressure, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"render", nullptr, Render, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"version", nullptr, Version, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof d / sizeof *d, d);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
