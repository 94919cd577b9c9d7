// ORACLE — test infrastructure only. Nothing in the shipped library may include, link or
// call this file; tests, smoke() and bench.py's CPU baseline use it as the checker.
//
// CPU restatement of Fire Pro's tile directory (src/engine/TileDirectory.ts) and of the
// pool-layout arithmetic (src/engine/types.ts:65-129), written sequentially and in double
// precision like the JavaScript it follows.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace oracle {

constexpr int TILE_BRICKS = 4;            // TileDirectory.ts:4
constexpr int TILE_BRICK_COUNT = 64;      // TileDirectory.ts:6
constexpr int LINGER_STEPS = 60;          // TileDirectory.ts:8
constexpr int TILE_FREE = 0, TILE_LIVE = 1, TILE_CLOSING = 2;  // TileDirectory.ts:10-12
constexpr int TILE_RECORD_WORDS = 32;     // TileDirectory.ts:15
constexpr int HASH_ROW = 256;             // TileDirectory.ts:17

// types.ts:56-60 — the WebGPU default limits the pools are sized for.
constexpr double MAX_TEXTURE_3D = 2048, MAX_BINDING = 134217728, MAX_BUFFER = 268435456;

// types.ts:87-92
inline int poolSlotsPerAxis(int brick) { return (int)std::floor(MAX_TEXTURE_3D / (brick + 1)); }

// types.ts:65-83
inline double maxPoolVoxels(int divisor, int brick) {
  double cells = std::pow((double)brick, 3);
  double bytes = std::max(std::pow((double)brick / divisor, 3) * 2, cells / 8);
  int perAxis = poolSlotsPerAxis(brick);
  double side = divisor == 1 ? std::min((double)perAxis, std::floor(MAX_TEXTURE_3D / (brick * 2)))
                             : (double)perAxis;
  return std::min(std::floor(std::min(MAX_BINDING, MAX_BUFFER) / bytes), side * side * perAxis) *
         cells;
}

// types.ts:100-113
inline int poolSide(int divisor, int brick) {
  int perAxis = poolSlotsPerAxis(brick);
  double slots = std::floor(maxPoolVoxels(divisor, brick) / std::pow((double)brick, 3));
  int donorSide =
      divisor == 1 ? (int)std::floor(MAX_TEXTURE_3D / (brick * 2)) : perAxis;
  return std::min(std::min(perAxis, donorSide),
                  std::max(4, (int)std::ceil(std::sqrt(slots / perAxis))));
}

// types.ts:116-119
inline int poolRows(int slots, int side) {
  int rows = std::max(1, (int)std::ceil((double)slots / side));
  return rows <= side ? rows : (int)std::ceil((double)rows / side) * side;
}

// types.ts:127-129
inline int initialPoolSlots(int side, int limit) {
  return std::min(limit, side * (int)std::ceil(128.0 / side));
}

// TileDirectory.ts:50-52, tiles.wgsl:52-56: 32-bit wrapping products.
inline uint32_t tileHash(int x, int y, int z) {
  return ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
}

struct Tile {
  int c[3];
  int slot;
  int64_t needed;
  bool closing;
};

struct TileBox {
  int lo[3], hi[3];
};

struct Tables {
  std::vector<int32_t> records;  // capacity * 32
  std::vector<int32_t> hash;     // size * 4
  std::vector<TileBox> clusters;
  std::vector<int32_t> tiles;    // capacity * 4
};

inline int64_t tileKey(int x, int y, int z) {
  return ((int64_t)(x & 0x1fffff) << 42) | ((int64_t)(y & 0x1fffff) << 21) | (int64_t)(z & 0x1fffff);
}

// TileDirectory (TileDirectory.ts:61-246). JavaScript Map iteration follows insertion
// order, which decides the hash placement and the cluster numbering, so `order` keeps it.
struct Directory {
  std::vector<Tile *> order;
  std::unordered_map<int64_t, Tile *> map;
  std::vector<Tile *> slots;
  bool dirty = true;
  int capacity = 0;
  std::vector<int> created;
  std::vector<std::array<int, 4>> snapshot;  // xyz + valid
  bool ground = true;
  int version = 0;

  ~Directory() { clear(); }

  int count() const { return (int)map.size(); }

  Tile *find(int x, int y, int z) const {
    auto it = map.find(tileKey(x, y, z));
    return it == map.end() ? nullptr : it->second;
  }

  // TileDirectory.ts:86-108
  void need(const int lo[3], const int hi[3], int64_t step) {
    for (int z = lo[2]; z <= hi[2]; z++)
      for (int y = ground ? std::max(0, lo[1]) : lo[1]; y <= hi[1]; y++)
        for (int x = lo[0]; x <= hi[0]; x++) {
          Tile *t = find(x, y, z);
          if (t) {
            t->needed = step;
            if (t->closing) {
              t->closing = false;
              dirty = true;
            }
            continue;
          }
          int slot = -1;
          for (size_t i = 0; i < slots.size(); i++)
            if (!slots[i]) {
              slot = (int)i;
              break;
            }
          if (slot < 0) slot = (int)slots.size();
          Tile *n = new Tile{{x, y, z}, slot, step, false};
          if (slot == (int)slots.size()) slots.push_back(n);
          else slots[slot] = n;
          map[tileKey(x, y, z)] = n;
          order.push_back(n);
          created.push_back(slot);
          dirty = true;
        }
  }

  // TileDirectory.ts:111-123
  void retire(int64_t step) {
    std::vector<Tile *> kept;
    kept.reserve(order.size());
    for (Tile *t : order) {
      if (t->closing) {
        map.erase(tileKey(t->c[0], t->c[1], t->c[2]));
        slots[t->slot] = nullptr;
        dirty = true;
        delete t;
        continue;
      }
      if (step - t->needed > LINGER_STEPS) {
        t->closing = true;
        dirty = true;
      }
      kept.push_back(t);
    }
    order.swap(kept);
    while (!slots.empty() && !slots.back()) slots.pop_back();
  }

  void clear() {
    for (Tile *t : order) delete t;
    order.clear();
    map.clear();
    slots.clear();
    created.clear();
    dirty = true;
  }

  int slotAt(int x, int y, int z) const {
    Tile *t = find(x, y, z);
    return t ? t->slot : -1;
  }

  int slotsUsed() const { return (int)slots.size(); }

  // TileDirectory.ts:140-144
  std::vector<int> takeCreated() {
    std::vector<int> out;
    for (int s : created)
      if (s < (int)slots.size() && slots[s]) out.push_back(s);
    created.clear();
    return out;
  }

  // TileDirectory.ts:151-186
  Tables tables(int cap) {
    if (cap < (int)slots.size()) throw std::runtime_error("Tile capacity is below the tiles in use.");
    capacity = cap;
    Tables t;
    t.records.assign((size_t)cap * TILE_RECORD_WORDS, -1);
    int size = HASH_ROW;
    while (size < (int)map.size() * 2) size *= 2;
    t.hash.assign((size_t)size * 4, -1);
    for (int s = 0; s < cap; s++) {
      int32_t *r = &t.records[(size_t)s * TILE_RECORD_WORDS];
      r[0] = r[1] = r[2] = 0;
      r[3] = TILE_FREE;
    }
    for (Tile *tile : order) {
      int x = tile->c[0], y = tile->c[1], z = tile->c[2];
      int32_t *r = &t.records[(size_t)tile->slot * TILE_RECORD_WORDS];
      r[0] = x;
      r[1] = y;
      r[2] = z;
      r[3] = tile->closing ? TILE_CLOSING : TILE_LIVE;
      uint32_t entry = tileHash(x, y, z) & (uint32_t)(size - 1);
      while (t.hash[entry * 4 + 3] >= 0) entry = (entry + 1) & (uint32_t)(size - 1);
      t.hash[entry * 4 + 0] = x;
      t.hash[entry * 4 + 1] = y;
      t.hash[entry * 4 + 2] = z;
      t.hash[entry * 4 + 3] = tile->slot;
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            Tile *n = find(x + dx, y + dy, z + dz);
            if (n) r[4 + (dz + 1) * 9 + (dy + 1) * 3 + dx + 1] = n->slot;
          }
    }
    std::unordered_map<const Tile *, int> clusterOf;
    clusters(t.clusters, clusterOf);
    t.tiles.assign((size_t)cap * 4, 0);
    for (int s = 0; s < cap; s++) {
      Tile *tile = s < (int)slots.size() ? slots[s] : nullptr;
      int32_t *w = &t.tiles[(size_t)s * 4];
      if (tile) {
        w[0] = tile->c[0];
        w[1] = tile->c[1];
        w[2] = tile->c[2];
        w[3] = clusterOf[tile];
      } else {
        w[0] = w[1] = w[2] = 0;
        w[3] = -1;
      }
    }
    snapshot.assign(slots.size(), {0, 0, 0, 0});
    for (size_t s = 0; s < slots.size(); s++)
      if (slots[s]) snapshot[s] = {slots[s]->c[0], slots[s]->c[1], slots[s]->c[2], 1};
    dirty = false;
    version++;
    return t;
  }

  // TileDirectory.ts:189-245
  void clusters(std::vector<TileBox> &out, std::unordered_map<const Tile *, int> &clusterOf) {
    std::unordered_map<const Tile *, int> group;
    std::vector<TileBox> boxes;
    for (Tile *first : order) {
      if (group.count(first)) continue;
      int index = (int)boxes.size();
      TileBox box{{first->c[0], first->c[1], first->c[2]}, {first->c[0], first->c[1], first->c[2]}};
      group[first] = index;
      std::vector<Tile *> stack{first};
      while (!stack.empty()) {
        Tile *t = stack.back();
        stack.pop_back();
        for (int a = 0; a < 3; a++) {
          box.lo[a] = std::min(box.lo[a], t->c[a]);
          box.hi[a] = std::max(box.hi[a], t->c[a]);
        }
        for (int dz = -1; dz <= 1; dz++)
          for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              Tile *n = find(t->c[0] + dx, t->c[1] + dy, t->c[2] + dz);
              if (n && !group.count(n)) {
                group[n] = index;
                stack.push_back(n);
              }
            }
      }
      boxes.push_back(box);
    }
    std::vector<int> into(boxes.size());
    for (size_t i = 0; i < into.size(); i++) into[i] = (int)i;
    std::function<int(int)> root = [&](int i) { return into[i] == i ? i : (into[i] = root(into[i])); };
    auto overlaps = [](const TileBox &a, const TileBox &b) {
      for (int k = 0; k < 3; k++)
        if (!(a.lo[k] <= b.hi[k] && b.lo[k] <= a.hi[k])) return false;
      return true;
    };
    for (bool merged = true; merged;) {
      merged = false;
      for (size_t i = 0; i < boxes.size(); i++)
        for (size_t j = i + 1; j < boxes.size(); j++) {
          int a = root((int)i), b = root((int)j);
          if (a == b || !overlaps(boxes[a], boxes[b])) continue;
          for (int k = 0; k < 3; k++) {
            boxes[a].lo[k] = std::min(boxes[a].lo[k], boxes[b].lo[k]);
            boxes[a].hi[k] = std::max(boxes[a].hi[k], boxes[b].hi[k]);
          }
          into[b] = a;
          merged = true;
        }
    }
    out.clear();
    std::unordered_map<int, int> renumbered;
    for (size_t i = 0; i < boxes.size(); i++)
      if (root((int)i) == (int)i) {
        renumbered[(int)i] = (int)out.size();
        out.push_back(boxes[i]);
      }
    for (auto &kv : group) clusterOf[kv.first] = renumbered[root(kv.second)];
  }
};

}  // namespace oracle
