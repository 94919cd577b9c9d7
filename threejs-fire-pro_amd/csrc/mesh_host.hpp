// Mesh emission on the host (MeshSourceRasterizer.ts:56-168, FluidSimulation.ts:1318-1352):
// the field cells that mesh sources' surface samples cover, the emission each covered cell
// receives, and per brick the table from its cells to those records.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "../../include/firepro.h"

namespace fp {
namespace mesh {

constexpr size_t MAX_SAMPLE_VISITS = 2000000;  // MeshSourceRasterizer.ts:24

// Cells of the world lattice, in the order they are first seen.
template <class T> struct CellMap {
  std::unordered_map<uint64_t, size_t> index;
  std::vector<std::array<int, 3>> cells;
  std::vector<T> values;
  static uint64_t key(int x, int y, int z) {
    return ((uint64_t)(uint32_t)(x + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(y + (1 << 20)) << 21) |
           (uint64_t)(uint32_t)(z + (1 << 20));
  }
  T &at(int x, int y, int z, const T &init) {
    auto r = index.emplace(key(x, y, z), values.size());
    if (r.second) {
      cells.push_back({x, y, z});
      values.push_back(init);
    }
    return values[r.first->second];
  }
};

// Emission of one covered cell, accumulated over the mesh sources.
struct Emission {
  double flame = 0, heat = 0, smoke = 0, fuel = 0, motion[3] = {0, 0, 0}, response = 0;
};

// The records of a frame's mesh sources: 8 floats each, (flame, heat rate, smoke rate, fuel
// rate) then (motion x, y, z, response), behind an all-zero record 0; and each record's cell.
struct Records {
  std::vector<float> data{std::vector<float>(8, 0.0f)};
  std::vector<std::array<int, 3>> cells;
  int count() const { return (int)cells.size(); }
};

// A mesh source's coverage of cells `spacing` wide (MeshSourceRasterizer.ts:98-121): each
// sample spreads weight * area over its eight trilinear cells, in units of the mean face
// area of a cell.
inline CellMap<double> coverage(const fp_source &s, double spacing) {
  if (!s.samples || s.sample_count <= 0 || !(s.surface_area > 0))
    throw std::invalid_argument("Mesh sources require surface samples and a positive surface area.");
  const double meanFace = (spacing * spacing + spacing * spacing + spacing * spacing) / 3;
  CellMap<double> coverage;
  for (int k = 0; k < s.sample_count; k++) {
    const double *p = s.samples + 4 * (size_t)k;
    int base[3];
    double frac[3];
    for (int a = 0; a < 3; a++) {
      const double g = p[a] / spacing - 0.5;
      base[a] = (int)std::floor(g);
      frac[a] = g - std::floor(g);
    }
    for (int z = 0; z < 2; z++)
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 2; x++) {
          const double tri = (x ? frac[0] : 1 - frac[0]) * (y ? frac[1] : 1 - frac[1]) * (z ? frac[2] : 1 - frac[2]);
          coverage.at(base[0] + x, base[1] + y, base[2] + z, 0.0) += p[3] * s.surface_area * tri / meanFace;
        }
  }
  return coverage;
}

// prepare (MeshSourceRasterizer.ts:66-166) with cells `spacing` wide.
inline Records rasterize(const fp_frame &f, double spacing) {
  CellMap<Emission> cells;
  size_t visits = 0;
  for (int i = 0; i < f.source_count; i++) {
    const fp_source &s = f.sources[i];
    if (s.shape != FP_SHAPE_MESH) continue;
    visits += (size_t)std::max(0, s.sample_count) * 8;
    if (visits > MAX_SAMPLE_VISITS) throw std::invalid_argument("Mesh source sampling exceeds two million cell visits.");
    const CellMap<double> coverage = mesh::coverage(s, spacing);
    for (size_t c = 0; c < coverage.values.size(); c++) {
      const double mask = std::min(1.0, coverage.values[c]);
      if (mask <= 0) continue;
      const auto &xyz = coverage.cells[c];
      Emission &e = cells.at(xyz[0], xyz[1], xyz[2], Emission());
      e.flame = std::max(e.flame, s.flame * mask);
      e.heat += s.heat_rate * mask;
      e.smoke += s.smoke_rate * mask;
      e.fuel += s.fuel_rate * mask;
      const double response = s.velocity_response * mask;
      for (int a = 0; a < 3; a++) e.motion[a] += s.velocity[a] * response;
      e.response += response;
    }
  }
  Records r;
  r.cells = cells.cells;
  r.data.reserve(8 * (cells.values.size() + 1));
  for (const Emission &e : cells.values) {
    const double v[8] = {e.flame, e.heat, e.smoke, e.fuel, e.motion[0], e.motion[1], e.motion[2], e.response};
    for (double x : v) r.data.push_back((float)x);
  }
  return r;
}

// The brick tables (FluidSimulation.ts:1329-1351): `bricks` words, each brick's table index
// or -1, then per table one record index per field cell (0 for none). `brickId` maps brick
// coordinates to a brick id, or -1 outside every tile.
template <class BrickId>
std::vector<int32_t> tables(const Records &r, int bricks, int shift, BrickId brickId) {
  const int edge = 1 << shift, mask = edge - 1, cells = edge * edge * edge;
  std::unordered_map<int, int> table;
  std::vector<std::pair<int64_t, int>> placed;
  for (int k = 0; k < r.count(); k++) {
    const auto &c = r.cells[k];
    const int id = brickId(c[0] >> shift, c[1] >> shift, c[2] >> shift);
    if (id < 0) continue;
    auto t = table.emplace(id, (int)table.size()).first->second;
    placed.push_back({(int64_t)t * cells + ((c[2] & mask) * edge + (c[1] & mask)) * edge + (c[0] & mask), k + 1});
  }
  std::vector<int32_t> out((size_t)bricks + table.size() * (size_t)cells, 0);
  std::fill(out.begin(), out.begin() + bricks, -1);
  for (const auto &t : table) out[t.first] = t.second;
  for (const auto &p : placed) out[(size_t)bricks + (size_t)p.first] = p.second;
  return out;
}

// The force coverage table (ForceFields.ts:114-179): per brick id of the velocity grid the
// first word of its entries or NO_COVERAGE; then per such brick a count and its entries, a
// force's index and one bit per velocity cell its mesh targets cover (coverage above zero).
constexpr uint32_t NO_COVERAGE = 0xffffffffu;
template <class BrickId>
std::vector<uint32_t> forceCoverage(const fp_frame &f, double spacing, int bricks, int shift, BrickId brickId) {
  const int side = 1 << shift, mask = side - 1;
  const int words = (side * side * side + 31) / 32;
  std::vector<int> order;  // brick ids in the order first seen
  std::unordered_map<int, std::vector<std::pair<int, std::vector<uint32_t>>>> entries;
  for (int i = 0; i < f.force_count; i++) {
    const fp_force &fc = f.forces[i];
    if (!fc.targeted) continue;
    for (int k = 0; k < fc.source_count; k++) {
      const fp_source &s = fc.sources[k];
      if (s.shape != FP_SHAPE_MESH) continue;
      const CellMap<double> cov = coverage(s, spacing);
      for (size_t c = 0; c < cov.values.size(); c++) {
        if (!(std::min(1.0, cov.values[c]) > 0)) continue;
        const auto &q = cov.cells[c];
        const int id = brickId(q[0] >> shift, q[1] >> shift, q[2] >> shift);
        if (id < 0) continue;
        auto found = entries.find(id);
        if (found == entries.end()) {
          order.push_back(id);
          found = entries.emplace(id, std::vector<std::pair<int, std::vector<uint32_t>>>()).first;
        }
        auto &list = found->second;
        std::vector<uint32_t> *bits = nullptr;
        for (auto &e : list)
          if (e.first == i) bits = &e.second;
        if (!bits) {
          list.push_back({i, std::vector<uint32_t>(words, 0u)});
          bits = &list.back().second;
        }
        const int bit = ((q[2] & mask) * side + (q[1] & mask)) * side + (q[0] & mask);
        (*bits)[bit >> 5] |= 1u << (bit & 31);
      }
    }
  }
  std::vector<uint32_t> out((size_t)bricks, NO_COVERAGE);
  for (int id : order) {
    out[id] = (uint32_t)out.size();
    const auto &list = entries[id];
    out.push_back((uint32_t)list.size());
    for (const auto &e : list) {
      out.push_back((uint32_t)e.first);
      out.insert(out.end(), e.second.begin(), e.second.end());
    }
  }
  return out;
}

}  // namespace mesh
}  // namespace fp
