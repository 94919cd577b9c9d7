// Host side of the render: the boxes the renderer marches and the blackbody table.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

#include "host_directory.hpp"

namespace fp {
namespace render {

// Bricks from lo to hi, inclusive (FluidSimulation.ts:121-124).
struct BrickBox {
  int lo[3], hi[3];
};

inline int floorDiv(int a, int b) {
  int q = a / b;
  return (a % b != 0 && (a < 0) != (b < 0)) ? q - 1 : q;
}

// The renderer's boxes (FluidSimulation.ts:913-969): one per cluster of tiles holding
// content at the latest measurement or a source's bricks, the union of those grown by
// BOX_MARGIN bricks inside the cluster's tiles, clusters ascending. `tiles` is every tile
// slot's coordinates with the index of its cluster's box, or -1.
struct Boxes {
  std::vector<BrickBox> boxes;
  std::vector<int> clusters;
  std::vector<int32_t> tiles;
  int version = 0;
  bool stale = true;  // the tile tables changed since the boxes were last published

  void update(const std::vector<std::pair<std::array<int, 3>, uint32_t>> &content, const std::vector<BrickBox> &extra,
              bool ground, const host::TileDirectory &dir, const std::vector<int32_t> &clusterTiles,
              const std::vector<host::TileBox> &tileClusters) {
    constexpr int TILE = 4, MARGIN = 2;
    std::vector<int> seen;
    std::vector<BrickBox> found;
    auto include = [&](const BrickBox &b) {
      int c[3];
      for (int a = 0; a < 3; a++) c[a] = floorDiv(b.lo[a] + b.hi[a], 2);
      if (ground) c[1] = std::max(0, c[1]);
      int slot = dir.slotAt(floorDiv(c[0], TILE), floorDiv(c[1], TILE), floorDiv(c[2], TILE));
      if (slot < 0) return;
      int cluster = clusterTiles[(size_t)slot * 4 + 3];
      auto it = std::find(seen.begin(), seen.end(), cluster);
      if (it == seen.end()) {
        seen.push_back(cluster);
        found.push_back(b);
        return;
      }
      BrickBox &u = found[it - seen.begin()];
      for (int a = 0; a < 3; a++) {
        u.lo[a] = std::min(u.lo[a], b.lo[a]);
        u.hi[a] = std::max(u.hi[a], b.hi[a]);
      }
    };
    // A content tile's activity word holds, per axis, the brick columns with content
    // (4 bits each); its box spans the lowest to the highest.
    for (const auto &ct : content) {
      BrickBox b;
      bool ok = true;
      for (int a = 0; a < 3; a++) {
        uint32_t used = (ct.second >> (4 * a)) & 15u;
        if (!used) {
          ok = false;
          break;
        }
        int lowest = __builtin_ctz(used), highest = 31 - __builtin_clz(used);
        b.lo[a] = ct.first[a] * TILE + lowest;
        b.hi[a] = ct.first[a] * TILE + highest;
      }
      if (ok) include(b);
    }
    for (const auto &b : extra) include(b);
    std::vector<int> order = seen;
    std::sort(order.begin(), order.end());
    std::vector<BrickBox> next;
    for (int cluster : order) {
      const BrickBox &u = found[std::find(seen.begin(), seen.end(), cluster) - seen.begin()];
      const host::TileBox &t = tileClusters[cluster];
      BrickBox b;
      for (int a = 0; a < 3; a++) {
        b.lo[a] = std::max(u.lo[a] - MARGIN, t.lo[a] * TILE);
        b.hi[a] = std::min(u.hi[a] + MARGIN, (t.hi[a] + 1) * TILE - 1);
      }
      next.push_back(b);
    }
    bool same = order == clusters && next.size() == boxes.size();
    for (size_t i = 0; same && i < next.size(); i++)
      for (int a = 0; a < 3; a++)
        if (next[i].lo[a] != boxes[i].lo[a] || next[i].hi[a] != boxes[i].hi[a]) same = false;
    if (same && !stale) return;
    stale = false;
    boxes = next;
    clusters = order;
    tiles = clusterTiles;
    for (size_t w = 3; w < tiles.size(); w += 4)
      if (tiles[w] >= 0) {
        auto it = std::find(order.begin(), order.end(), tiles[w]);
        tiles[w] = it == order.end() ? -1 : (int32_t)(it - order.begin());
      }
    version++;
  }
};

// The blackbody table (blackbody.ts:5-97): 512 texels from 500 K to 10000 K, each the
// peak-normalized linear RGB of Planck radiance integrated against the Wyman-Sloan-Shirley
// multi-lobe CIE 1931 observer fit, and log2 of its peak relative to 1800 K's.
inline std::vector<float> blackbody() {
  constexpr int N = 512;
  struct Lobe {
    double height, center, left, right;
  };
  static const Lobe X[] = {{0.362, 442, 0.0624, 0.0374}, {1.056, 599.8, 0.0264, 0.0323}, {-0.065, 501.1, 0.049, 0.0382}};
  static const Lobe Y[] = {{0.821, 568.8, 0.0213, 0.0247}, {0.286, 530.9, 0.0613, 0.0322}};
  static const Lobe Z[] = {{1.217, 437, 0.0845, 0.0278}, {0.681, 459, 0.0385, 0.0725}};
  auto fit = [](const Lobe *lobes, int n, double wl) {
    double sum = 0;
    for (int k = 0; k < n; k++) {
      double d = (wl - lobes[k].center) * (wl < lobes[k].center ? lobes[k].left : lobes[k].right);
      sum = sum + lobes[k].height * std::exp(-0.5 * d * d);
    }
    return sum;
  };
  std::vector<std::array<double, 3>> observer(95);
  for (int i = 0; i < 95; i++) {
    double wl = 360 + i * 5;
    observer[i] = {fit(X, 3, wl), fit(Y, 2, wl), fit(Z, 2, wl)};
  }
  // Planck's law in W sr^-1 m^-3 with the exact SI constants.
  auto radiance = [](double nm, double kelvin) {
    const double h = 6.62607015e-34, c = 299792458, k = 1.380649e-23;
    double wl = nm * 1e-9;
    return (2 * h * c * c) / (std::pow(wl, 5) * std::expm1((h * c) / (wl * k * kelvin)));
  };
  auto linearRgb = [&](double kelvin, double out[3]) {
    double xyz[3] = {0, 0, 0};
    for (int i = 0; i < 95; i++) {
      double e = radiance(360 + i * 5, kelvin) * 5e-9;
      for (int c = 0; c < 3; c++) xyz[c] += observer[i][c] * e;
    }
    out[0] = std::max(0.0, 3.2406 * xyz[0] - 1.5372 * xyz[1] - 0.4986 * xyz[2]);
    out[1] = std::max(0.0, -0.9689 * xyz[0] + 1.8758 * xyz[1] + 0.0415 * xyz[2]);
    out[2] = std::max(0.0, 0.0557 * xyz[0] - 0.204 * xyz[1] + 1.057 * xyz[2]);
  };
  double ref[3];
  linearRgb(1800, ref);
  const double refPeak = std::max({ref[0], ref[1], ref[2]});
  std::vector<float> table((size_t)N * 4);
  for (int i = 0; i < N; i++) {
    double rgb[3];
    linearRgb(500 + ((10000 - 500) * (double)i) / (N - 1), rgb);
    double peak = std::max({rgb[0], rgb[1], rgb[2]});
    float *o = &table[(size_t)i * 4];
    if (peak <= 0) {
      o[0] = o[1] = o[2] = 0.0f;
      o[3] = -32.0f;
      continue;
    }
    for (int c = 0; c < 3; c++) o[c] = (float)(rgb[c] / peak);
    o[3] = (float)std::max(-32.0, std::log2(peak / refPeak));
  }
  return table;
}

}  // namespace render
}  // namespace fp
