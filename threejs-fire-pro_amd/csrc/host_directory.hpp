// Host side of the sparse grid: the tile directory (src/engine/TileDirectory.ts:61-246) and
// the pool-layout arithmetic (src/engine/types.ts:65-129). The directory keeps its tiles
// in creation order, because that order decides where tiles land in the open-addressed
// hash and how clusters are numbered, both of which the GPU tables reproduce exactly.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <climits>
#include <utility>
#include <vector>

namespace fp {
namespace host {

constexpr double MAX_TEXTURE_3D = 2048, MAX_BINDING = 134217728, MAX_BUFFER = 268435456;  // types.ts:56-60

inline int poolSlotsPerAxis(int brick) { return (int)std::floor(MAX_TEXTURE_3D / (brick + 1)); }
inline double maxPoolVoxels(int divisor, int brick) {
  double cells = (double)brick * brick * brick;
  double vb = (double)brick / divisor;
  double bytes = std::max(vb * vb * vb * 2, cells / 8);
  double perAxis = poolSlotsPerAxis(brick);
  double side = divisor == 1 ? std::min(perAxis, std::floor(MAX_TEXTURE_3D / (brick * 2))) : perAxis;
  return std::min(std::floor(std::min(MAX_BINDING, MAX_BUFFER) / bytes), side * side * perAxis) * cells;
}
inline int poolSide(int divisor, int brick) {
  int perAxis = poolSlotsPerAxis(brick);
  double slots = std::floor(maxPoolVoxels(divisor, brick) / ((double)brick * brick * brick));
  int donorSide = divisor == 1 ? (int)std::floor(MAX_TEXTURE_3D / (brick * 2)) : perAxis;
  return std::min(std::min(perAxis, donorSide), std::max(4, (int)std::ceil(std::sqrt(slots / perAxis))));
}
inline int poolRows(int slots, int side) {
  int rows = std::max(1, (int)std::ceil((double)slots / side));
  return rows <= side ? rows : (int)std::ceil((double)rows / side) * side;
}
inline int initialPoolSlots(int side, int limit) { return std::min(limit, side * (int)std::ceil(128.0 / side)); }

struct TileBox {
  int lo[3], hi[3];
};

// Tile coordinates -> tile, open-addressed with linear probing, at most half full.
template <class V> class FlatIndex {
 public:
  static constexpr int64_t EMPTY = INT64_MIN;
  size_t size() const { return count_; }
  V find(int64_t k) const {
    if (keys_.empty()) return V();
    for (size_t h = mix(k) & mask_;; h = (h + 1) & mask_) {
      if (keys_[h] == k) return vals_[h];
      if (keys_[h] == EMPTY) return V();
    }
  }
  void insert(int64_t k, V v) {
    if ((count_ + 1) * 2 > keys_.size()) grow();
    size_t h = mix(k) & mask_;
    while (keys_[h] != EMPTY && keys_[h] != k) h = (h + 1) & mask_;
    if (keys_[h] == EMPTY) count_++;
    keys_[h] = k;
    vals_[h] = v;
  }
  void clear() {
    keys_.clear();
    vals_.clear();
    count_ = 0;
    mask_ = 0;
  }
  // Rebuild from (key, value) pairs, after removals.
  template <class It> void rebuild(It begin, It end) {
    clear();
    for (It i = begin; i != end; ++i) insert(i->first, i->second);
  }

 private:
  static size_t mix(int64_t k) {
    uint64_t x = (uint64_t)k * 0x9E3779B97F4A7C15ull;
    return (size_t)(x ^ (x >> 29));
  }
  void grow() {
    std::vector<int64_t> keys(std::max<size_t>(64, keys_.size() * 2), EMPTY);
    std::vector<V> vals(keys.size());
    std::swap(keys, keys_);
    std::swap(vals, vals_);
    mask_ = keys_.size() - 1;
    count_ = 0;
    for (size_t i = 0; i < keys.size(); i++)
      if (keys[i] != EMPTY) insert(keys[i], vals[i]);
  }
  std::vector<int64_t> keys_;
  std::vector<V> vals_;
  size_t count_ = 0, mask_ = 0;
};

class TileDirectory {
 public:
  struct Tile {
    int c[3];
    int slot;
    int64_t needed;
    bool closing;
  };
  static constexpr int LINGER_STEPS = 60, RECORD_WORDS = 32, HASH_ROW = 256;

  bool ground = true;
  bool dirty = true;
  int version = 0;
  int capacity = 0;

  ~TileDirectory() { clear(); }

  int count() const { return (int)index_.size(); }
  int slotsUsed() const { return (int)slots_.size(); }

  static int64_t key(int x, int y, int z) {
    return ((int64_t)(x & 0x1fffff) << 42) | ((int64_t)(y & 0x1fffff) << 21) | (int64_t)(z & 0x1fffff);
  }
  const Tile *find(int x, int y, int z) const { return index_.find(key(x, y, z)); }
  int slotAt(int x, int y, int z) const {
    const Tile *t = find(x, y, z);
    return t ? t->slot : -1;
  }

  // need (TileDirectory.ts:86-108): refresh or create every tile of the box.
  void need(const int lo[3], const int hi[3], int64_t step) {
    for (int z = lo[2]; z <= hi[2]; z++)
      for (int y = ground ? std::max(0, lo[1]) : lo[1]; y <= hi[1]; y++)
        for (int x = lo[0]; x <= hi[0]; x++) {
          Tile *found = index_.find(key(x, y, z));
          if (found) {
            found->needed = step;
            if (found->closing) {
              found->closing = false;
              dirty = true;
            }
            continue;
          }
          // The lowest free slot (TileDirectory.ts:100); none below firstFree_.
          int slot = firstFree_;
          while (slot < (int)slots_.size() && slots_[slot]) slot++;
          Tile *t = new Tile{{x, y, z}, slot, step, false};
          if (slot == (int)slots_.size()) slots_.push_back(t);
          else slots_[slot] = t;
          firstFree_ = slot + 1;
          index_.insert(key(x, y, z), t);
          order_.push_back(t);
          created_.push_back(slot);
          dirty = true;
        }
  }

  // retire (TileDirectory.ts:111-123).
  void retire(int64_t step) {
    size_t w = 0;
    bool removed = false;
    for (size_t r = 0; r < order_.size(); r++) {
      Tile *t = order_[r];
      if (t->closing) {
        slots_[t->slot] = nullptr;
        firstFree_ = std::min(firstFree_, t->slot);
        dirty = true;
        removed = true;
        delete t;
        continue;
      }
      if (step - t->needed > LINGER_STEPS) {
        t->closing = true;
        dirty = true;
      }
      order_[w++] = t;
    }
    order_.resize(w);
    while (!slots_.empty() && !slots_.back()) slots_.pop_back();
    firstFree_ = std::min(firstFree_, (int)slots_.size());
    if (removed) {
      std::vector<std::pair<int64_t, Tile *>> live;
      live.reserve(order_.size());
      for (Tile *t : order_) live.push_back({key(t->c[0], t->c[1], t->c[2]), t});
      index_.rebuild(live.begin(), live.end());
    }
  }

  void clear() {
    for (Tile *t : order_) delete t;
    order_.clear();
    index_.clear();
    slots_.clear();
    firstFree_ = 0;
    created_.clear();
    dirty = true;
  }

  std::vector<int> takeCreated() {
    std::vector<int> out;
    for (int s : created_)
      if (s < (int)slots_.size() && slots_[s]) out.push_back(s);
    created_.clear();
    return out;
  }

  // tables (TileDirectory.ts:151-186): `words` gets the cap tile records with 27 neighbor
  // slots, RECORD_WORDS each, followed by the hash (hashWords words, 4 a row); `tiles` the
  // per-slot (coordinates, cluster) words; `coords` snapshots the slots' tiles.
  void tables(int cap, std::vector<int32_t> &words, int &hashWords, std::vector<TileBox> &clusters,
              std::vector<int32_t> &tiles) {
    if (cap < (int)slots_.size()) throw std::runtime_error("Tile capacity is below the tiles in use.");
    capacity = cap;
    int size = HASH_ROW;
    while (size < (int)index_.size() * 2) size *= 2;
    hashWords = size * 4;
    const size_t recordWords = (size_t)cap * RECORD_WORDS;
    words.resize(recordWords + (size_t)hashWords);
    // Free slots: coordinates and the live word zero, no neighbors.
    for (int s = 0; s < cap; s++) {
      int32_t *r = &words[(size_t)s * RECORD_WORDS];
      r[0] = r[1] = r[2] = r[3] = 0;
      std::fill(r + 4, r + RECORD_WORDS, -1);
    }
    int32_t *const hash = words.data() + recordWords;
    std::fill(hash, hash + hashWords, -1);
    int32_t *const records = words.data();
    const uint32_t mask = (uint32_t)size - 1;
    buildDense();
    for (const Tile *t : order_) {
      int32_t *r = records + (size_t)t->slot * RECORD_WORDS;
      r[0] = t->c[0];
      r[1] = t->c[1];
      r[2] = t->c[2];
      r[3] = t->closing ? 2 : 1;
      uint32_t h = (((uint32_t)t->c[0] * 73856093u) ^ ((uint32_t)t->c[1] * 19349663u) ^
                    ((uint32_t)t->c[2] * 83492791u)) & mask;
      while (hash[(size_t)h * 4 + 3] >= 0) h = (h + 1) & mask;
      for (int k = 0; k < 3; k++) hash[(size_t)h * 4 + k] = t->c[k];
      hash[(size_t)h * 4 + 3] = t->slot;
      neighborSlots(t, r + 4);
    }
    std::vector<int> clusterOf;  // per slot
    buildClusters(clusters, clusterOf);
    tiles.assign((size_t)cap * 4, 0);
    coords.assign(slots_.size(), {0, 0, 0, 0});
    for (int s = 0; s < cap; s++) {
      const Tile *t = s < (int)slots_.size() ? slots_[s] : nullptr;
      int32_t *w = &tiles[(size_t)s * 4];
      if (t) {
        w[0] = t->c[0];
        w[1] = t->c[1];
        w[2] = t->c[2];
        w[3] = clusterOf[s];
        coords[s] = {t->c[0], t->c[1], t->c[2], 1};
      } else {
        w[3] = -1;
      }
    }
    dirty = false;
    version++;
  }

  std::vector<std::array<int, 4>> coords;
  const std::vector<Tile *> &order() const { return order_; }

 private:
  // clusters (TileDirectory.ts:189-245): 26-connected groups, boxed, merged until disjoint.
  // clusterOf maps each slot to its cluster (-1 for a free slot).
  void buildClusters(std::vector<TileBox> &out, std::vector<int> &clusterOf) {
    std::vector<int> &group = group_;  // per slot: the connected group
    group.assign(slots_.size(), -1);
    std::vector<TileBox> boxes;
    std::vector<int> &stack = stack_;
    for (const Tile *first : order_) {
      if (group[first->slot] >= 0) continue;
      int idx = (int)boxes.size();
      TileBox box{{first->c[0], first->c[1], first->c[2]}, {first->c[0], first->c[1], first->c[2]}};
      group[first->slot] = idx;
      stack.assign(1, first->slot);
      while (!stack.empty()) {
        const Tile *t = slots_[stack.back()];
        stack.pop_back();
        for (int a = 0; a < 3; a++) {
          box.lo[a] = std::min(box.lo[a], t->c[a]);
          box.hi[a] = std::max(box.hi[a], t->c[a]);
        }
        int32_t nb[27];
        neighborSlots(t, nb);
        for (int k = 0; k < 27; k++) {
          const int n = nb[k];
          if (n >= 0 && group[n] < 0) {
            group[n] = idx;
            stack.push_back(n);
          }
        }
      }
      boxes.push_back(box);
    }
    std::vector<int> into(boxes.size());
    for (size_t i = 0; i < into.size(); i++) into[i] = (int)i;
    auto root = [&](int i) {
      int r = i;
      while (into[r] != r) r = into[r];
      while (into[i] != r) {
        int n = into[i];
        into[i] = r;
        i = n;
      }
      return r;
    };
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
    std::vector<int> renumbered(boxes.size(), -1);
    for (size_t i = 0; i < boxes.size(); i++)
      if (root((int)i) == (int)i) {
        renumbered[i] = (int)out.size();
        out.push_back(boxes[i]);
      }
    clusterOf.assign(slots_.size(), -1);
    for (size_t s = 0; s < slots_.size(); s++)
      if (group[s] >= 0) clusterOf[s] = renumbered[root(group[s])];
  }

  // A dense grid of the tiles' slots (-1 where none) over their bounding box with one free
  // cell on every side, for the neighbor lookups of tables(); the flat index serves when the
  // box is too large.
  void buildDense() {
    dense_.clear();
    if (order_.empty()) return;
    int lo[3], hi[3];
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = order_[0]->c[a];
    for (const Tile *t : order_)
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], t->c[a]);
        hi[a] = std::max(hi[a], t->c[a]);
      }
    size_t cells = 1;
    for (int a = 0; a < 3; a++) {
      denseLo_[a] = lo[a] - 1;
      denseDim_[a] = hi[a] - lo[a] + 3;
      cells *= (size_t)denseDim_[a];
    }
    if (cells > ((size_t)1 << 24)) return;
    dense_.assign(cells, -1);
    for (const Tile *t : order_) dense_[denseCell(t)] = t->slot;
    const ptrdiff_t sy = denseDim_[0], sz = (ptrdiff_t)denseDim_[0] * denseDim_[1];
    for (int k = 0; k < 27; k++) denseOff_[k] = (k / 9 - 1) * sz + ((k / 3) % 3 - 1) * sy + (k % 3 - 1);
  }
  size_t denseCell(const Tile *t) const {
    return ((size_t)(t->c[2] - denseLo_[2]) * denseDim_[1] + (t->c[1] - denseLo_[1])) * denseDim_[0] +
           (t->c[0] - denseLo_[0]);
  }
  // The slots of the 27 tiles around tile t (itself included), (dz * 3 + dy) * 3 + dx order,
  // -1 where none. Inside the dense grid's free border a tile's neighbors are fixed offsets
  // from its cell, with no bounds tests.
  void neighborSlots(const Tile *t, int32_t out[27]) const {
    if (dense_.empty()) {
      for (int k = 0; k < 27; k++) {
        const Tile *n = index_.find(key(t->c[0] + k % 3 - 1, t->c[1] + (k / 3) % 3 - 1, t->c[2] + k / 9 - 1));
        out[k] = n ? n->slot : -1;
      }
      return;
    }
    const int32_t *base = &dense_[denseCell(t)];
    for (int k = 0; k < 27; k++) out[k] = base[denseOff_[k]];
  }

  std::vector<Tile *> order_;
  FlatIndex<Tile *> index_;
  int firstFree_ = 0;
  std::vector<int32_t> dense_;
  int denseLo_[3] = {0, 0, 0}, denseDim_[3] = {0, 0, 0};
  ptrdiff_t denseOff_[27] = {};
  std::vector<int> group_, stack_;
  std::vector<Tile *> slots_;
  std::vector<int> created_;
};

}  // namespace host
}  // namespace fp
