// The solver runtime and the C ABI (include/firepro.h): host orchestration of one step
// (src/engine/FluidSimulation.ts substep, :1374-1531) over device-resident pools.
//
// One HIP stream per simulation. The step enqueues allocation, transport, forces, the
// multigrid V-cycle and the projection, then copies the activity readback (active count,
// free slots, failures, per-tile content) into pinned memory and waits for it: the next
// step's tile directory sees this step's content, the earliest a WebGPU mapAsync can
// deliver it (FluidSimulation.ts:1701-1746).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/firepro.h"
#include "host_directory.hpp"
#include "kernels.hpp"
#include "render_host.hpp"
#include "mesh_host.hpp"

namespace fp {

thread_local std::string g_error;

struct InvalidArgument : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DeviceFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHECK(expr)                                                                              \
  do {                                                                                              \
    hipError_t err_ = (expr);                                                                       \
    if (err_ != hipSuccess)                                                                         \
      throw DeviceFailure(std::string(#expr) + ": " + hipGetErrorString(err_));                     \
  } while (0)

// ---- recycled device and pinned host blocks --------------------------------------------
// Simulations come and go while others run (a serving loop creates and destroys one per
// request), and hipMalloc / hipFree / hipHostMalloc cost tens of microseconds to
// milliseconds each. Freed blocks are kept per device in size classes (quarter steps
// between powers of two, so at most 25% over the request) and handed out again. A block
// freed while work enqueued before the free may still touch it waits in `pending`; only a
// device-wide synchronization moves pending blocks to `ready`, so a block is never reused
// before every earlier user of it has finished, whatever stream that user ran on. The cache
// holds at most `limit` bytes; beyond that, and when an allocation fails, blocks go back to
// the runtime.
class BlockCache {
 public:
  BlockCache(hipError_t (*get)(void **, size_t), hipError_t (*put)(void *), size_t limit)
      : get_(get), put_(put), limit_(limit) {}

  static size_t sizeClass(size_t bytes) {
    if (bytes <= 256) return 256;
    int k = 63 - __builtin_clzll((unsigned long long)bytes);
    size_t step = (size_t)1 << (k - 2);
    return (bytes + step - 1) / step * step;
  }

  void *take(size_t bytes) {
    const size_t c = sizeClass(bytes);
    {
      std::lock_guard<std::mutex> lock(mu_);
      Pool &pool = pools_[device()];
      auto it = pool.ready.find(c);
      if (it == pool.ready.end() && pool.pendingClasses.count(c)) {
        HIPCHECK(hipDeviceSynchronize());
        settle(pool);
        it = pool.ready.find(c);
      }
      if (it != pool.ready.end()) {
        void *p = it->second;
        pool.ready.erase(it);
        pool.bytes -= c;
        return p;
      }
    }
    void *p = nullptr;
    if (get_(&p, c) != hipSuccess) {
      (void)hipGetLastError();
      trim();
      HIPCHECK(get_(&p, c));
    }
    return p;
  }

  void give(void *p, size_t bytes) {
    if (!p) return;
    const size_t c = sizeClass(bytes);
    std::lock_guard<std::mutex> lock(mu_);
    Pool &pool = pools_[device()];
    if (pool.bytes + c > limit_) {
      (void)put_(p);  // the runtime's free waits for the device itself
      return;
    }
    pool.pending.push_back({c, p});
    pool.pendingClasses.insert(c);
    pool.bytes += c;
  }

  // Returns every cached block of the current device to the runtime.
  void trim() {
    std::lock_guard<std::mutex> lock(mu_);
    Pool &pool = pools_[device()];
    (void)hipDeviceSynchronize();
    settle(pool);
    for (auto &kv : pool.ready) (void)put_(kv.second);
    pool.ready.clear();
    pool.bytes = 0;
  }

 private:
  struct Pool {
    std::multimap<size_t, void *> ready;
    std::vector<std::pair<size_t, void *>> pending;
    std::multiset<size_t> pendingClasses;
    size_t bytes = 0;
  };
  static int device() {
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
  }
  static void settle(Pool &pool) {
    for (auto &b : pool.pending) pool.ready.insert(b);
    pool.pending.clear();
    pool.pendingClasses.clear();
  }
  hipError_t (*get_)(void **, size_t);
  hipError_t (*put_)(void *);
  size_t limit_;
  std::mutex mu_;
  std::map<int, Pool> pools_;
};

static hipError_t deviceGet(void **p, size_t bytes) { return hipMalloc(p, bytes); }
static hipError_t deviceFree(void *p) { return hipFree(p); }
static hipError_t pinnedGet(void **p, size_t bytes) { return hipHostMalloc(p, bytes, hipHostMallocDefault); }
static hipError_t pinnedFree(void *p) { return hipHostFree(p); }

// Device blocks, up to 64 GiB cached; pinned host blocks, up to 1 GiB.
BlockCache &deviceBlocks() {
  static BlockCache cache(deviceGet, deviceFree, (size_t)64 << 30);
  return cache;
}
BlockCache &pinnedBlocks() {
  static BlockCache cache(pinnedGet, pinnedFree, (size_t)1 << 30);
  return cache;
}
template <class T> T *pinnedAlloc(size_t count) { return static_cast<T *>(pinnedBlocks().take(count * sizeof(T))); }
template <class T> void pinnedRelease(T *&p, size_t count) {
  pinnedBlocks().give(p, count * sizeof(T));
  p = nullptr;
}

// Streams of finished simulations, idle, for the next ones.
class StreamPool {
 public:
  hipStream_t take() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      int d = 0;
      (void)hipGetDevice(&d);
      auto &v = free_[d];
      if (!v.empty()) {
        hipStream_t s = v.back();
        v.pop_back();
        return s;
      }
    }
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
  }
  // `s` has no work left.
  void give(hipStream_t s) {
    if (!s) return;
    std::lock_guard<std::mutex> lock(mu_);
    int d = 0;
    (void)hipGetDevice(&d);
    free_[d].push_back(s);
  }

 private:
  std::mutex mu_;
  std::map<int, std::vector<hipStream_t>> free_;
};
StreamPool &streams() {
  static StreamPool pool;
  return pool;
}

template <class T> struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) p = static_cast<T *>(deviceBlocks().take(count * sizeof(T)));
  }
  // Reallocate to `count`, keeping the first min(n, count) elements.
  void grow(size_t count, hipStream_t s) {
    if (count <= n) return;
    T *q = static_cast<T *>(deviceBlocks().take(count * sizeof(T)));
    if (p && n) HIPCHECK(hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    release();
    p = q;
    n = count;
  }
  void release() {
    if (p) deviceBlocks().give(p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// ---- small housekeeping kernels ---------------------------------------------------------

// growFreeList (fluid-pool-growth.wgsl:12-21): new slots go under the stack.
__global__ void k_grow_free_list(uint32_t *freeSlots, int *poolState, int first, int added) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < added) freeSlots[i] = (uint32_t)(first + added - 1 - i);
  if (i == 0) poolState[0] += added;
}
// A created tile's bricks start without slots (FluidSimulation.ts:728-734).
__global__ void k_clear_tiles(int *pages, int *entries, uint32_t *state, uint32_t *activity, const int *slots,
                              int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count * 64) return;
  int id = slots[i >> 6] * 64 + (i & 63);
  pages[id] = -1;
  entries[id] = -1;
  state[id] = 0u;
  activity[id] = 0u;
}
__global__ void k_fill_int(int *p, int value, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}
__global__ void k_fill_u32(uint32_t *p, uint32_t value, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}
// Failed allocations count from one measurement to the next (FluidSimulation.ts:1525): the
// lane that reads the counter restarts it.
__global__ void k_readback(const int *counts, int *poolState, const uint32_t *tileFlags, int cap, int *out) {
  readbackLane(counts, poolState, tileFlags, cap, out, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}

// The readbacks of a batch, a workgroup per simulation, each into its own words of one
// buffer; the failure counters restart for the next step.
struct ReadbackArgs {
  const int *counts;
  int *poolState;
  const uint32_t *tileFlags;
  int cap;
  int *out;
  const double *norms;  // defect correction's norms, or null
  double *report;       // their relative residual and cycle count
};
// One simulation's activity words and defect-correction report; entries past the batch's
// members are empty (no output).
__device__ __forceinline__ void readbackOne(const ReadbackArgs &a) {
  if (!a.out) return;
  if (threadIdx.x == 0) {
    a.out[0] = a.counts[0];
    a.out[1] = a.poolState[0];
    a.out[2] = a.poolState[1];
    a.out[3] = a.counts[1] + a.counts[5];
    a.poolState[1] = 0;
    if (a.norms) {
      a.report[0] = a.norms[4];
      a.report[1] = a.norms[6];
    }
  }
  for (int t = threadIdx.x; t < a.cap; t += blockDim.x) a.out[4 + t] = (int)a.tileFlags[t];
}
__global__ void __launch_bounds__(256) k_readback_batch(const ReadbackArgs *args, int n) {
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    readbackOne(args[i]);
    __syncthreads();
  }
}

// ---- the curl-noise table (curl-noise.wgsl:1-45), built once on the host -----------------

static float noiseHash(float px, float py, float pz) {
  uint32_t v[3] = {(uint32_t)(int)px & 15u, (uint32_t)(int)py & 15u, (uint32_t)(int)pz & 15u};
  for (uint32_t &c : v) c = c * 1664525u + 1013904223u;
  v[0] += v[1] * v[2];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
  for (uint32_t &c : v) c ^= c >> 16;
  v[0] += v[1] * v[2];
  v[1] += v[2] * v[0];
  v[2] += v[0] * v[1];
  return (float)(v[0] >> 8) / 16777216.0f;
}
static float hmix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
static void noiseGradient(float px, float py, float pz, float out[3]) {
  float ix = std::floor(px), iy = std::floor(py), iz = std::floor(pz);
  float f[3] = {px - ix, py - iy, pz - iz};
  float s[3], ds[3];
  for (int a = 0; a < 3; a++) {
    s[a] = f[a] * f[a] * f[a] * (f[a] * (f[a] * 6.0f - 15.0f) + 10.0f);
    ds[a] = 30.0f * f[a] * f[a] * (f[a] - 1.0f) * (f[a] - 1.0f);
  }
  float a = noiseHash(ix, iy, iz), b = noiseHash(ix + 1, iy, iz), c = noiseHash(ix, iy + 1, iz);
  float d = noiseHash(ix + 1, iy + 1, iz), e = noiseHash(ix, iy, iz + 1), f1 = noiseHash(ix + 1, iy, iz + 1);
  float g = noiseHash(ix, iy + 1, iz + 1), h = noiseHash(ix + 1, iy + 1, iz + 1);
  out[0] = hmix(hmix(b - a, d - c, s[1]), hmix(f1 - e, h - g, s[1]), s[2]) * ds[0];
  out[1] = hmix(hmix(c - a, d - b, s[0]), hmix(g - e, h - f1, s[0]), s[2]) * ds[1];
  out[2] = hmix(hmix(e - a, f1 - b, s[0]), hmix(g - c, h - d, s[0]), s[1]) * ds[2];
}
static std::vector<float4> curlNoiseTable() {
  std::vector<float4> t(64 * 64 * 64);
  for (int z = 0; z < 64; z++)
    for (int y = 0; y < 64; y++)
      for (int x = 0; x < 64; x++) {
        float p[3] = {((float)x + 0.5f) / 64.0f * 16.0f, ((float)y + 0.5f) / 64.0f * 16.0f,
                      ((float)z + 0.5f) / 64.0f * 16.0f};
        float dx[3], dy[3], dz[3];
        noiseGradient(p[0], p[1], p[2], dx);
        noiseGradient(p[0] + 7.1f, p[1] + 3.7f, p[2] + 11.3f, dy);
        noiseGradient(p[0] + 5.2f, p[1] + 7.9f, p[2] + 9.5f, dz);
        t[((size_t)z * 64 + y) * 64 + x] = make_float4(dz[1] - dy[2], dx[2] - dz[0], dy[0] - dx[1], 0.0f);
      }
  return t;
}

// The table on the current device, built and uploaded at its first use and shared,
// read-only, by every simulation on that device until the process ends.
static const float4 *sharedNoise() {
  static std::mutex mu;
  static std::map<int, float4 *> tables;
  std::lock_guard<std::mutex> lock(mu);
  int d = 0;
  HIPCHECK(hipGetDevice(&d));
  float4 *&t = tables[d];
  if (!t) {
    static const std::vector<float4> host = curlNoiseTable();
    float4 *p = nullptr;
    HIPCHECK(hipMalloc(&p, host.size() * sizeof(float4)));
    HIPCHECK(hipMemcpy(p, host.data(), host.size() * sizeof(float4), hipMemcpyHostToDevice));
    t = p;
  }
  return t;
}

// ---- the allocation-free scene lights (fp_sim_scene_lights_noalloc) ----
// Their buffers live in one file-scope device array, `scratch`, at fixed offsets: the
// anchors, the blackbody table, per-brick totals and nearest anchors, per-block totals and
// the anchors' totals. Each GPU holds its own copy until the process ends.
constexpr int NOALLOC_MAX_BRICKS = 65536;
constexpr int NOALLOC_MAX_ANCHORS = 64;
constexpr size_t NA_ANCHORS = 0;                                                  // float4 per anchor
constexpr size_t NA_LUT = NA_ANCHORS + (size_t)NOALLOC_MAX_ANCHORS * sizeof(float4);  // 512 float4
constexpr size_t NA_PARTIALS = NA_LUT + 512 * sizeof(float4);                     // 8 floats per brick
constexpr size_t NA_BRICK_ANCHOR = NA_PARTIALS + (size_t)NOALLOC_MAX_BRICKS * 8 * sizeof(float);
constexpr size_t NA_BLOCK_SUMS = NA_BRICK_ANCHOR + (size_t)NOALLOC_MAX_BRICKS * sizeof(int);
constexpr size_t NA_SUMS =
    NA_BLOCK_SUMS + (size_t)(NOALLOC_MAX_BRICKS / LIGHT_BLOCK) * NOALLOC_MAX_ANCHORS * 8 * sizeof(float);
constexpr size_t NA_BYTES = NA_SUMS + (size_t)NOALLOC_MAX_ANCHORS * 8 * sizeof(float);
static_assert(NA_LUT % 256 == 0 && NA_PARTIALS % 256 == 0 && NA_BRICK_ANCHOR % 256 == 0 &&
                  NA_BLOCK_SUMS % 256 == 0 && NA_SUMS % 256 == 0,
              "scratch regions start on 256-byte boundaries");

}  // namespace fp

// In host code this name is a stand-in, not a device address: host code uses the address
// hipGetSymbolAddress gives for the current device (fp::noallocScratch).
static __device__ __align__(256) unsigned char scratch[fp::NA_BYTES];

namespace fp {

// The device address of `scratch` on the current device, found once per device (with that
// device current) and kept; the first call also writes the blackbody table into it.
static unsigned char *noallocScratch() {
  static std::mutex mu;
  static std::map<int, unsigned char *> addresses;
  std::lock_guard<std::mutex> lock(mu);
  int d = 0;
  HIPCHECK(hipGetDevice(&d));
  unsigned char *&ptr = addresses[d];
  if (!ptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw DeviceFailure(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    static const std::vector<float> lut = render::blackbody();  // host, built once per process
    HIPCHECK(hipMemcpy(static_cast<unsigned char *>(p) + NA_LUT, lut.data(), lut.size() * sizeof(float),
                       hipMemcpyHostToDevice));
    ptr = static_cast<unsigned char *>(p);
  }
  return ptr;
}

constexpr int MIN_TILES = 16;
// Workgroups that walk a step's freshly allocated bricks (a few per step, rarely thousands).
constexpr unsigned FRESH_GRID = 2048;
constexpr double MAX_DT = 1.0 / 30.0;

}  // namespace fp

using namespace fp;

struct fp_sim {
  // Grid (construction-only) and derived layout.
  double voxel = 0;
  int d = 1, B = 16, b = 16, sF = 4, sV = 4, nF = 0, nV = 0, side = 0, maxSlots = 0, brickLevels = 0, bottom = 0;
  // Coarse smoke: the divisor, the smoke brick's edge, shift and cells, and the smoke cells
  // a velocity cell spans on each axis (FluidSimulation.ts:374-380).
  int smokeDiv = 1, smokeEdge = 0, smokeShift = 0, nS = 0, smokeRatio = 1;
  bool maccormack = true;
  // Settings written before each step.
  double voxelBudget = INFINITY, cutoff = 0;
  bool ground = true;
  // Host state (FluidSimulation.ts:189-355).
  int poolSlots = 0, usedSlots = 0, poolTarget = 0;
  int64_t poolGrownAt = 0;
  bool poolLimited = false;
  int64_t activeBricks = 0;
  double time = 0;
  int64_t steps = 0;
  bool mayHaveFlame = false, solidsStale = true, geometryDirty = true;
  // The tile set changed since the topology was last built (and nothing else that
  // geometryDirty stands for): tile slots moved, bricks came and went.
  bool tilesMoved = false;
  hipEvent_t readbackEvent = nullptr;  // recorded after each step's activity readback
  // Work an entry point left queued on the stream when it returned (a step's tail after its
  // readback, a render into a caller's device buffer); drain() waits for it.
  bool tailPending = false;
  int colliderCount = 0;
  std::vector<uint32_t> prevColliders, prevMasks;
  std::vector<uint64_t> sourceIds;
  std::vector<std::pair<std::array<int, 3>, uint32_t>> contentTiles;
  int freeCount = 0, lastFailures = 0, lastZeroed = 0;
  host::TileDirectory dir;
  int capacity = 0;
  std::vector<int32_t> tableWords, tilesHost;
  std::vector<int> created;  // tile slots created by the last table upload
  std::vector<host::TileBox> clusters;
  uint32_t hashMask = 0;
  hipStream_t stream = nullptr;
  bool stepped = false;

  // Per brick id.
  DBuf<int> pages, entries, records, active, zeroList, poolState, counts, readback;
  // The tile tables: capacity records of 32 words, then the tile hash from word hashOffset,
  // then the slots the last upload created.
  DBuf<int> tileRec;
  size_t hashOffset = 0;
  DBuf<int> tileLists;  // the tile lists of movedTileTopology
  std::vector<int> tileListsHost;
  DBuf<uint32_t> brickState, activity, stateTmp, rankTmp, blockSums, blockPrefix, freeSlots;
  // Pools, per slot.
  DBuf<float4> F, Fp, Fn, Curl;
  DBuf<float4> Don;  // the scalar backtraces of a step (advection to correction)
  DBuf<float4> Vtr;  // the velocity backtraces of a step
  const float4 *noise = nullptr;  // the shared curl-noise table
  DBuf<float> E, V, Vadv, Vcor, Vf, D, P0, P1;
  DBuf<float> S, Sp, Sn;  // coarse smoke density, predictor and next (smokeDivisor 2 or 4)
  // Mesh emission: the records (8 floats each behind a zero record) and the brick tables,
  // with what they were built from.
  DBuf<float> meshRecs;
  DBuf<int> meshTables;
  mesh::Records meshHost;
  std::vector<int32_t> meshTablesHost;
  std::vector<double> meshKey;
  int meshTablesVersion = -1, meshBricks = -1, tilesVersion = 0;
  DBuf<uint32_t> solids;
  struct LevelBufs {
    DBuf<float> p, q, rhs;
    DBuf<uint8_t> topo, eflag;
    int side = 0, tile = 0;
  };
  LevelBufs lv[MAX_LEVELS];
  float *latestPressure = nullptr;
  float *latestPressureLo = nullptr;  // the low part when the pressure is a pair
  // showSolverField: the field copied after every step (FP_SOLVER_FIELD_*) and its copy,
  // a velocity-grid float per cell in slot order.
  int solverField = 0;
  DBuf<float> solverView;
  // Pressure policy: 0 the reference's one cold V-cycle; 1-3 defect correction (an FP32
  // outer loop to a relative residual) around V-cycles whose levels are stored in FP32,
  // BF16 or FP16, with the coarsest level solved in FP64.
  int pressureMode = 0;
  double pressureTol = 1e-6;
  int pressureMaxCycles = 20;
  int lastCycles = 1;
  double lastResidual = -1.0;
  struct HalfLevel {
    DBuf<uint16_t> p, q, rhs;
  };
  HalfLevel lv16[MAX_LEVELS];
  DBuf<float> dcRhs, dcX, dcXlo, dcR;
  DBuf<double> dcPartials, dcNorms;
  double *pinnedNorms = nullptr;
  // Render: the boxes (updated with the tiles each step) and the preparation's state
  // carried between renders (VolumeLighting.ts:28-63).
  render::Boxes rboxes;
  struct RenderCache {
    int version = -1;
    bool stale = true;
    int64_t lastStep = -1, lastLightStep = -1, occupancyStep = -1;
    std::string lastKey;
    int poolSlots = -1, lightDivisor = -1, lastWork = -1;
    int boxCount = 0, pageCount = 0;
    std::vector<int32_t> boxWords;
    DBuf<int4> boxes, tiles;
    DBuf<int> table;
    DBuf<float4> R, L, lut, image;
    DBuf<float> RS;  // smoothed coarse smoke with aprons (smokeDivisor 2 or 4)
    DBuf<float4> slices;  // the half-resolution march's depth slices
    // The slice march's 16 x 16 tiles: last frame's cost per tile and the order it implies
    // (longest first); orderTiles is the tile count the order was made for (0: none yet).
    DBuf<uint32_t> tileCost;
    // The latest render's sample counts (fp_sim_render_stats): the slices' positions visited
    // and samples shaded, then the march's.
    DBuf<unsigned long long> sampleCounts;
    DBuf<int> tileOrder;
    int orderTiles = 0;
    DBuf<uint8_t> O;
    // Scene lights: anchors, per-brick totals and anchors, per-block totals, the totals.
    DBuf<float4> anchors;
    DBuf<float> partials, blockSums, sums;
    DBuf<int> brickAnchor;
  } rc;
  // Scene.
  DBuf<Emitter> emitters;
  DBuf<Collider> colliders;
  DBuf<uint32_t> colliderMasks;
  DBuf<uint32_t> forceCoverage;  // velocity cells mesh targets cover, by brick (ForceFields.ts)
  std::vector<uint32_t> coverageHost;
  bool hasForceCoverage = false;
  DBuf<ForceRec> forces;
  DBuf<Region> regions;
  int emitterCount = 0, forceCount = 0;
  Params P{};
  // Readback.
  int *pinned = nullptr;
  size_t pinnedWords = 0;
  // Profiling.
  bool profiling = false;
  int profileLevel = 0;
  using Clock = std::chrono::steady_clock;
  // A host phase of the step, in ms.
  void hostMark(const char *name, Clock::time_point t0) {
    if (!profiling) return;
    double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    auto &e = profile[name];
    if (e.second == 0 && std::find(profileNames.begin(), profileNames.end(), name) == profileNames.end())
      profileNames.push_back(name);
    e.first += ms;
    e.second += 1;
  }
  struct Mark {
    const char *name;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> eventPool;
  std::map<std::string, std::pair<double, int>> profile;
  std::vector<std::string> profileNames;

  int brickCount() const { return capacity * 64; }
  int slotLimit() const {
    double voxels = std::min(voxelBudget, host::maxPoolVoxels(d, B));
    return std::max(1, (int)std::min((double)maxSlots, std::floor(voxels / ((double)B * B * B))));
  }
  double tileSize() const { return 4.0 * B * voxel; }
  double velocitySpacing() const { return voxel * d; }

  Layout layout() const {
    Layout L;
    L.B = B;
    L.b = b;
    L.sF = sF;
    L.sV = sV;
    L.nF = nF;
    L.nV = nV;
    L.side = side;
    L.brickLevels = brickLevels;
    L.sd = smokeDiv;
    L.bs = smokeEdge;
    L.sS = smokeShift;
    L.nS = nS;
    return L;
  }
  Dir dirArgs() const {
    Dir r;
    r.tileRec = tileRec.p;
    r.hash = tileRec.p ? tileRec.p + hashOffset : nullptr;
    r.hashMask = hashMask;
    r.entries = entries.p;
    r.records = records.p;
    r.capacity = capacity;
    return r;
  }

  // ---- profiling marks ----
  hipEvent_t event() {
    if (!eventPool.empty()) {
      hipEvent_t e = eventPool.back();
      eventPool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  // A device interval; at level 2 only the dominant stage's.
  int begin(bool dominant = false) {
    if (!profiling || (profileLevel == 2 && !dominant)) return -1;
    Mark m{nullptr, event(), event()};
    HIPCHECK(hipEventRecord(m.a, stream));
    marks.push_back(m);
    return (int)marks.size() - 1;
  }
  void end(int i, const char *name) {
    if (i < 0) return;
    marks[i].name = name;
    HIPCHECK(hipEventRecord(marks[i].b, stream));
  }
  void collect() {
    for (Mark &m : marks) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, m.a, m.b));
      auto &e = profile[m.name];
      if (e.second == 0 && std::find(profileNames.begin(), profileNames.end(), m.name) == profileNames.end())
        profileNames.push_back(m.name);
      e.first += ms;
      e.second += 1;
      eventPool.push_back(m.a);
      eventPool.push_back(m.b);
    }
    marks.clear();
  }

  // ---- construction and allocation (FluidSimulation.ts:359-532) ----
  void create(const fp_grid &g) {
    if (!std::isfinite(g.voxel_size) || g.voxel_size < 0.0001 || g.voxel_size > 16)
      throw InvalidArgument("voxelSize must be from 0.0001 to 16 meters.");
    voxel = g.voxel_size;
    d = g.velocity_divisor;
    if (d != 1 && d != 2 && d != 4) throw InvalidArgument("velocityDivisor must be 1, 2 or 4.");
    B = g.brick_size;
    if (B != 8 && B != 16 && B != 32) throw InvalidArgument("brickSize must be 8, 16 or 32 fine cells.");
    smokeDiv = g.smoke_divisor;
    if (smokeDiv != 1 && smokeDiv != 2 && smokeDiv != 4) throw InvalidArgument("smokeDivisor must be 1, 2 or 4.");
    maccormack = g.scalar_maccormack != 0;
    b = B / d;
    sF = 0;
    while ((1 << sF) < B) sF++;
    sV = 0;
    while ((1 << sV) < b) sV++;
    nF = B * B * B;
    nV = b * b * b;
    if (b < 2) throw InvalidArgument("velocity bricks need at least 2 cells a side.");
    smokeEdge = B / smokeDiv;
    smokeShift = 0;
    while ((1 << smokeShift) < smokeEdge) smokeShift++;
    nS = smokeEdge * smokeEdge * smokeEdge;
    smokeRatio = d > smokeDiv ? d / smokeDiv : 1;
    side = host::poolSide(d, B);
    maxSlots = side * side * host::poolSlotsPerAxis(B);
    brickLevels = sV;
    bottom = brickLevels + 2;
    if (bottom + 1 > MAX_LEVELS) throw InvalidArgument("too many multigrid levels.");
    stream = streams().take();
    poolState.alloc(4);
    counts.alloc(8);
    noise = sharedNoise();
    allocate();
  }

  void allocate() {
    drain();  // the last step's queued tail finishes before its buffers are replaced
    int slots = host::initialPoolSlots(side, slotLimit());
    poolSlots = 0;
    usedSlots = 0;
    poolTarget = 0;
    sourceIds.clear();
    poolLimited = false;
    allocatePools(slots, false);
    std::vector<uint32_t> stack(slots);
    for (int i = 0; i < slots; i++) stack[i] = (uint32_t)(slots - 1 - i);
    freeSlots.alloc(slots);
    HIPCHECK(hipMemcpyAsync(freeSlots.p, stack.data(), slots * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    int state[4] = {slots, 0, 0, 0};
    HIPCHECK(hipMemcpyAsync(poolState.p, state, sizeof(state), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    poolSlots = slots;
    freeCount = slots;
    solidsStale = true;
    geometryDirty = true;
    dir.clear();
    contentTiles.clear();
    capacity = 0;
    growTiles(MIN_TILES);
    applyTiles();
    HIPCHECK(hipStreamSynchronize(stream));  // no step's readback follows the uploads yet
    rboxes.update(contentTiles, {}, ground, dir, tilesHost, clusters);
  }

  // Pools for `slots` slots; with `keep`, live state (fields, expansion, velocity, solids)
  // carries over and every slot keeps its place.
  void allocatePools(int slots, bool keep) {
    drain();  // the last step's queued tail finishes before its buffers are replaced
    size_t s = (size_t)slots;
    if (keep) {
      F.grow(s * nF, stream);
      E.grow(s * nF, stream);
      V.grow(s * nV * 3, stream);
      if (smokeDiv > 1) S.grow(s * nS, stream);
      solids.grow(s * (nF / 32), stream);
    } else {
      F.alloc(s * nF);
      E.alloc(s * nF);
      V.alloc(s * nV * 3);
      if (smokeDiv > 1) S.alloc(s * nS);
      solids.alloc(s * (nF / 32));
      HIPCHECK(hipMemsetAsync(solids.p, 0, solids.bytes(), stream));
    }
    Fp.alloc(s * nF);
    if (smokeDiv > 1) {
      Sp.alloc(s * nS);
      Sn.alloc(s * nS);
    }
    Don.alloc(s * nF * 3);
    Vtr.alloc(s * nV * 4);
    Fn.alloc(s * nF);
    Curl.alloc(s * nV);
    Vadv.alloc(s * nV * 3);
    Vcor.alloc(s * nV * 3);
    Vf.alloc(s * nV * 3);
    D.alloc(s * nV);
    P0.alloc(s * nV);
    P1.alloc(s * nV);
    latestPressure = P0.p;
    latestPressureLo = nullptr;
    for (int L = 0; L < brickLevels; L++) {
      int sd = b >> L;
      size_t n = s * sd * sd * sd;
      lv[L].side = sd;
      lv[L].tile = 0;
      if (L > 0) {
        lv[L].p.alloc(n);
        lv[L].q.alloc(n);
        lv[L].rhs.alloc(n);
      }
      lv[L].topo.alloc(n);
      lv[L].eflag.alloc(n);
    }
  }

  // Tile resources for `cap` tile slots, keeping each brick's page, state, flags, entry and
  // record (FluidSimulation.ts:663-705).
  void growTiles(int cap) {
    drain();  // the last step's queued tail finishes before its buffers are replaced
    size_t bricks = (size_t)cap * 64;
    size_t old = (size_t)capacity * 64;
    auto growFill = [&](DBuf<int> &buf, size_t count, int fill) {
      buf.grow(count, stream);
      if (count > old) k_fill_int<<<(unsigned)((count - old + 255) / 256), 256, 0, stream>>>(buf.p + old, fill, count - old);
    };
    growFill(pages, bricks, -1);
    growFill(entries, bricks, -1);
    records.grow(bricks * RECORD_WORDS, stream);
    brickState.grow(bricks, stream);
    if (bricks > old) k_fill_u32<<<(unsigned)((bricks - old + 255) / 256), 256, 0, stream>>>(brickState.p + old, 0u, bricks - old);
    // Brick flags are kept; the per-tile words after them are rewritten every step.
    {
      DBuf<uint32_t> next;
      next.alloc(bricks + cap);
      k_fill_u32<<<(unsigned)((bricks + cap + 255) / 256), 256, 0, stream>>>(next.p, 0u, bricks + cap);
      if (activity.p && old)
        HIPCHECK(hipMemcpyAsync(next.p, activity.p, old * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      activity.release();
      activity = next;
      next.p = nullptr;
      next.n = 0;
    }
    stateTmp.alloc(bricks);
    rankTmp.alloc(bricks);
    size_t blocks = (bricks + 255) / 256;
    blockSums.alloc(blocks * 3);
    blockPrefix.alloc(blocks * 3);
    active.alloc(bricks);
    zeroList.alloc(bricks * 4);
    tileRec.alloc((size_t)cap * 32 + 4 * host::TileDirectory::HASH_ROW);
    hashOffset = (size_t)cap * 32;
    capacity = cap;
    for (int L = brickLevels; L <= bottom; L++) {
      int sd = 4 >> (L - brickLevels);
      size_t n = (size_t)cap * sd * sd * sd;
      lv[L].side = sd;
      lv[L].tile = 1;
      lv[L].p.alloc(n);
      lv[L].q.alloc(n);
      lv[L].rhs.alloc(n);
      lv[L].topo.alloc(n);
      lv[L].eflag.alloc(n);
    }
    size_t words = 4 + (size_t)cap;
    readback.alloc(words);
    if (words > pinnedWords) {
      if (pinned) pinnedRelease(pinned, pinnedWords);
      pinned = pinnedAlloc<int>(words);
      pinnedWords = words;
    }
    geometryDirty = true;
  }

  // applyTiles (FluidSimulation.ts:663-737).
  void applyTiles() {
    tilesVersion++;
    int needed = std::max(MIN_TILES, dir.slotsUsed());
    if (needed > capacity) {
      int cap = capacity;
      while (cap < needed) cap *= 2;
      growTiles(cap);
    }
    Clock::time_point t0 = Clock::now();
    int hashWords = 0;
    dir.tables(capacity, tableWords, hashWords, clusters, tilesHost);
    hostMark("host_tiles_tables", t0);
    // Records, hash and the created slots in one copy.
    const size_t tableSize = tableWords.size();
    created = dir.takeCreated();
    tableWords.insert(tableWords.end(), created.begin(), created.end());
    if (tileRec.n < tableWords.size()) {
      drain();  // the last step's queued tail may still read the old block
      tileRec.alloc(tableWords.size() + tableWords.size() / 4);
    }
    HIPCHECK(hipMemcpyAsync(tileRec.p, tableWords.data(), tableWords.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                            stream));
    hashOffset = tableSize - (size_t)hashWords;
    hashMask = (uint32_t)(hashWords / 4 - 1);
    if (!created.empty()) {
      int n = (int)created.size() * 64;
      k_clear_tiles<<<(n + 255) / 256, 256, 0, stream>>>(pages.p, entries.p, brickState.p, activity.p,
                                                          tileRec.p + tableSize, (int)created.size());
    }
    // The copy reads tableWords when the stream reaches it; each step waits for its
    // readback, queued after it, before the words are rebuilt.
    tilesMoved = true;
    rboxes.stale = true;  // FluidSimulation.ts:708-709
  }

  // growPool (FluidSimulation.ts:1013-1093).
  void growPool(int target) {
    int limit = slotLimit();
    int rows = host::poolRows(std::min(target, limit), side);
    int slots = std::min(limit, rows * side);
    if (slots <= poolSlots) return;
    int added = slots - poolSlots;
    allocatePools(slots, true);
    DBuf<uint32_t> next;
    next.alloc(slots);
    HIPCHECK(hipMemcpyAsync(next.p + added, freeSlots.p, (size_t)poolSlots * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                            stream));
    k_grow_free_list<<<(added + 255) / 256, 256, 0, stream>>>(next.p, poolState.p, poolSlots, added);
    HIPCHECK(hipStreamSynchronize(stream));
    freeSlots.release();
    freeSlots = next;
    next.p = nullptr;
    next.n = 0;
    freeCount += added;
    poolSlots = slots;
    poolGrownAt = steps;
    geometryDirty = true;
  }

  // ---- one step ----
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
    double margin = 2 * B * voxel + tileSize();
    std::vector<render::BrickBox> sourceBricks;
    for (int i = 0; i < f.source_count; i++) {
      const fp_source &s = f.sources[i];
      double lo[3], hi[3];
      if (s.shape == FP_SHAPE_MESH && s.samples && s.sample_count > 0) {
        // A mesh source needs the bounds of its samples, a cell past the margin.
        for (int a = 0; a < 3; a++) {
          lo[a] = INFINITY;
          hi[a] = -INFINITY;
        }
        for (int k = 0; k < s.sample_count; k++)
          for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], s.samples[4 * (size_t)k + a]);
            hi[a] = std::max(hi[a], s.samples[4 * (size_t)k + a]);
          }
        needBox(lo, hi, margin + voxel);
      } else {
        double extent = std::max(std::max(std::max(s.size[0], s.size[1]), s.size[2]), 0.03) + s.outward_speed * 0.25;
        for (int a = 0; a < 3; a++) {
          lo[a] = s.position[a] - extent;
          hi[a] = s.position[a] + extent;
        }
        needBox(lo, hi, margin);
      }
      render::BrickBox bb;  // bricksOver (FluidSimulation.ts:862-869)
      for (int a = 0; a < 3; a++) {
        bb.lo[a] = (int)std::floor(lo[a] / (B * voxel));
        bb.hi[a] = (int)std::floor(hi[a] / (B * voxel));
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
    if (dir.dirty) {
      Clock::time_point a0 = Clock::now();
      applyTiles();
      hostMark("host_tiles_apply", a0);
    }
    Clock::time_point b0 = Clock::now();
    rboxes.update(contentTiles, sourceBricks, ground, dir, tilesHost, clusters);
    hostMark("host_tiles_boxes", b0);
  }

  static double finite32(const char *name, double v) {
    float f = (float)v;
    if (!std::isfinite(v) || !std::isfinite(f)) throw InvalidArgument(std::string(name) + " must be finite in f32.");
    return (double)f;
  }
  // packFlameStep (flame.ts:43-87) with its validation.
  void packFlame(const fp_flame &fl, double dt) {
    double stepf = finite32("Flame timestep", dt), lifespan = finite32("Flame lifespan", fl.lifespan);
    if (stepf <= 0 || lifespan <= 0) throw InvalidArgument("Flame timestep and lifespan must be positive in f32.");
    finite32("Lifetime decrement", stepf / lifespan);
    double smoke = finite32("Smoke rate", fl.smoke_rate), heat = finite32("Heat rate", fl.heat_rate);
    double expansion = finite32("Expansion rate", fl.expansion_rate);
    if (smoke < 0 || heat < 0) throw InvalidArgument("Smoke and heat rates must be nonnegative.");
    for (double rate : {smoke, heat, expansion}) finite32("Integrated flame output", rate * lifespan);
    const char *names[3] = {"smoke", "heat", "expansion"};
    for (int ch = 0; ch < 3; ch++) {
      int n = fl.knot_count[ch];
      if (n < 2 || n > 8) throw InvalidArgument(std::string(names[ch]) + " ramp needs 2–8 knots.");
      double previous = -1;
      for (int i = 0; i < n; i++) {
        double x = finite32("lifetime", fl.knots[ch][i][0]), y = finite32("response", fl.knots[ch][i][1]);
        if (x < 0 || x > 1 || y < 0 || y > 1 || x <= previous)
          throw InvalidArgument(std::string(names[ch]) + " knots must increase in f32 with coordinates in [0,1].");
        if ((i == 0 && x != 0) || (i == n - 1 && x != 1))
          throw InvalidArgument(std::string(names[ch]) + " ramp must span lifetime 0–1.");
        P.knots[ch][i][0] = (float)x;
        P.knots[ch][i][1] = (float)y;
        previous = x;
      }
      P.knotCount[ch] = n;
    }
    P.fdt = (float)stepf;
    P.lifespan = (float)lifespan;
    P.fsmoke = (float)smoke;
    P.fheat = (float)heat;
    P.fexpansion = (float)expansion;
  }

  // A scene array's upload. Kernels only read these arrays, so a block that already holds
  // the same bytes (same block, same contents as its last upload) is left as it is: a static
  // scene costs no copy per step.
  struct Uploaded {
    const void *block = nullptr;
    std::vector<unsigned char> bytes;
  };
  std::map<const void *, Uploaded> uploaded;
  template <class T> void upload(DBuf<T> &buf, const std::vector<T> &host) {
    size_t n = std::max<size_t>(1, host.size());
    if (buf.n < n) {
      drain();  // the last step's queued tail may still read the old block
      buf.alloc(n + n / 2);
    }
    if (host.empty()) return;
    const size_t bytes = host.size() * sizeof(T);
    Uploaded &last = uploaded[&buf];
    const unsigned char *src = reinterpret_cast<const unsigned char *>(host.data());
    if (last.block == buf.p && last.bytes.size() == bytes && std::memcmp(last.bytes.data(), src, bytes) == 0) return;
    HIPCHECK(hipMemcpyAsync(buf.p, host.data(), bytes, hipMemcpyHostToDevice, stream));
    last.block = buf.p;
    last.bytes.assign(src, src + bytes);
  }

  // The box of field cells holding every cell the colliders make solid (Params::solidLo and
  // solidHi), with two cells of margin. A sphere or box collider's cells satisfy
  // |axes_k . (w - center)| < extent_k on every axis k (insideCollider), so w - center =
  // A^-1 proj lies within sum_k |A^-1_jk| extent_k on world axis j; a mesh's cells are its
  // box. Solid bits are always built from the current colliders (a change rebuilds them
  // all), so none is set outside. Singular axes give the unbounded box.
  int solidLo[3] = {0, 0, 0}, solidHi[3] = {-1, -1, -1};
  void solidBox(const std::vector<Collider> &cs) {
    const double big = 1e9;
    double lo[3] = {big, big, big}, hi[3] = {-big, -big, -big};
    bool unbounded = false;
    for (const Collider &c : cs) {
      if (c.box == 2) {
        for (int a = 0; a < 3; a++) {
          lo[a] = std::min(lo[a], (double)c.lo[a]);
          hi[a] = std::max(hi[a], (double)c.lo[a] + c.size[a] - 1);
        }
        continue;
      }
      double m[3][3];  // rows: the axes
      for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) m[k][a] = c.axes[k][a];
      const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                         m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                         m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
      if (!(std::fabs(det) > 1e-6)) {
        unbounded = true;
        continue;
      }
      double inv[3][3];  // inv[j][k] = (A^-1)_jk
      for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) {
          const int r0 = (k + 1) % 3, r1 = (k + 2) % 3, c0 = (j + 1) % 3, c1 = (j + 2) % 3;
          inv[j][k] = (m[r0][c0] * m[r1][c1] - m[r0][c1] * m[r1][c0]) / det;
        }
      for (int j = 0; j < 3; j++) {
        double h = 0;
        for (int k = 0; k < 3; k++) h += std::fabs(inv[j][k]) * std::fabs((double)c.extent[k]);
        lo[j] = std::min(lo[j], std::floor(((double)c.center[j] - h) / voxel - 0.5));
        hi[j] = std::max(hi[j], std::ceil(((double)c.center[j] + h) / voxel - 0.5));
      }
    }
    for (int a = 0; a < 3; a++) {
      const bool none = lo[a] > hi[a];
      solidLo[a] = unbounded ? INT32_MIN : none ? 0 : (int)std::max(lo[a] - 2, -2e9);
      solidHi[a] = unbounded ? INT32_MAX : none ? -1 : (int)std::min(hi[a] + 2, 2e9);
    }
  }

  // packScene (FluidSimulation.ts:1150-1366): sphere/box emitters and colliders, analytic
  // forces. Returns whether the colliders' packed words changed.
  std::vector<Emitter> hostEmitters;
  std::vector<Collider> hostColliders;
  std::vector<ForceRec> hostForces;
  std::vector<Region> hostRegions;
  // Mesh emission records and brick tables (FluidSimulation.ts:1318-1352), rebuilt when the
  // mesh sources or the tiles change.
  void packMesh(const fp_frame &f) {
    std::vector<double> key;
    for (int i = 0; i < f.source_count; i++) {
      const fp_source &s = f.sources[i];
      if (s.shape != FP_SHAPE_MESH) continue;
      key.insert(key.end(), {(double)s.sample_count, s.surface_area, s.flame, s.heat_rate, s.smoke_rate, s.fuel_rate,
                             s.velocity[0], s.velocity[1], s.velocity[2], s.velocity_response});
      if (s.samples && s.sample_count > 0) key.insert(key.end(), s.samples, s.samples + 4 * (size_t)s.sample_count);
    }
    bool records = key != meshKey;
    if (records) {
      try {
        meshHost = mesh::rasterize(f, voxel);
      } catch (const std::invalid_argument &e) {
        throw InvalidArgument(e.what());
      }
      meshKey = key;
      upload(meshRecs, meshHost.data);
    }
    if (meshHost.count() == 0) {
      meshTablesVersion = -1;
      return;
    }
    if (records || meshTablesVersion != tilesVersion || meshBricks != brickCount()) {
      meshTablesHost = mesh::tables(meshHost, brickCount(), sF, [&](int bx, int by, int bz) {
        int slot = dir.slotAt(bx >> 2, by >> 2, bz >> 2);
        return slot < 0 ? -1 : slot * 64 + ((bz & 3) * 4 + (by & 3)) * 4 + (bx & 3);
      });
      upload(meshTables, meshTablesHost);
      meshTablesVersion = tilesVersion;
      meshBricks = brickCount();
    }
  }

  bool packScene(const fp_frame &f) {
    packMesh(f);
    hostEmitters.clear();
    for (int i = 0; i < f.source_count; i++) {
      const fp_source &s = f.sources[i];
      if (s.shape == FP_SHAPE_MESH) continue;  // emits through the mesh tables
      if (s.shape != FP_SHAPE_SPHERE && s.shape != FP_SHAPE_BOX)
        throw InvalidArgument("source.shape must be sphere, box or mesh.");
      Emitter e;
      for (int a = 0; a < 3; a++) {
        e.pos[a] = (float)s.position[a];
        e.size[a] = (float)s.size[a];
        e.vel[a] = (float)s.velocity[a];
      }
      e.box = s.shape == FP_SHAPE_BOX;
      e.flame = (float)s.flame;
      e.heat = (float)s.heat_rate;
      e.smoke = (float)s.smoke_rate;
      e.response = (float)s.velocity_response;
      e.outward = (float)s.outward_speed;
      e.period = (float)(s.has_variation ? s.variation_period : 0.5);
      e.strength = (float)(s.has_variation ? s.variation_strength : 0.0);
      e.fuel = (float)s.fuel_rate;
      for (int a = 0; a < 3; a++) e.invSize[a] = 1.0f / std::max(e.size[a], 0.03f);
      e.invPeriod = 1.0f / e.period;
      hostEmitters.push_back(e);
    }
    int n = f.collider_count;
    std::vector<uint32_t> packed((size_t)std::max(1, n) * 28, 0u);
    std::vector<uint32_t> masks;
    hostColliders.clear();
    for (int i = 0; i < n; i++) {
      const fp_collider &c = f.colliders[i];
      if (c.shape == FP_SHAPE_MESH) {
        // A mesh's occupancy, packed one bit per cell (FluidSimulation.ts:1221-1260).
        if (!c.occupancy || c.cells[0] <= 0 || c.cells[1] <= 0 || c.cells[2] <= 0)
          throw InvalidArgument("Mesh collider occupancy does not match its size.");
        Collider cr{};
        cr.box = 2;
        cr.maskOffset = (int)masks.size();
        const size_t cells = (size_t)c.cells[0] * c.cells[1] * c.cells[2];
        masks.resize(masks.size() + (cells + 31) / 32, 0u);
        for (size_t k = 0; k < cells; k++)
          if (c.occupancy[k]) masks[cr.maskOffset + (k >> 5)] |= 1u << (k & 31);
        int32_t words[8] = {c.lo[0], c.lo[1], c.lo[2], cr.maskOffset, c.cells[0], c.cells[1], c.cells[2], 0};
        std::memcpy(&packed[(size_t)i * 28 + 20], words, sizeof(words));
        const float kind = 2.0f;
        std::memcpy(&packed[(size_t)i * 28 + 3], &kind, 4);
        for (int a = 0; a < 3; a++) {
          cr.lo[a] = c.lo[a];
          cr.size[a] = c.cells[a];
        }
        hostColliders.push_back(cr);
        continue;
      }
      if (c.shape != FP_SHAPE_SPHERE && c.shape != FP_SHAPE_BOX)
        throw InvalidArgument("collider.shape.type must be sphere, box or mesh.");
      float w[20] = {0};
      for (int a = 0; a < 3; a++) {
        w[a] = (float)c.position[a];
        w[4 + a] = (float)c.size[a];
        for (int k = 0; k < 3; k++) w[8 + 4 * k + a] = (float)c.axes[k][a];
      }
      w[3] = c.shape == FP_SHAPE_BOX ? 1.0f : 0.0f;
      std::memcpy(&packed[(size_t)i * 28], w, sizeof(w));
      Collider cr{};
      for (int a = 0; a < 3; a++) {
        cr.center[a] = w[a];
        cr.extent[a] = w[4 + a];
        for (int k = 0; k < 3; k++) cr.axes[k][a] = w[8 + 4 * k + a];
      }
      cr.box = c.shape == FP_SHAPE_BOX;
      hostColliders.push_back(cr);
    }
    solidBox(hostColliders);
    bool changed = packed != prevColliders || n != colliderCount || masks != prevMasks;
    prevColliders = packed;
    prevMasks = masks;
    colliderCount = n;
    upload(colliderMasks, masks);
    hostForces.clear();
    hostRegions.clear();
    bool anyMeshTargets = false;
    for (int i = 0; i < f.force_count; i++) {
      const fp_force &fc = f.forces[i];
      if (fc.type < 0 || fc.type > 3) throw InvalidArgument("force.type must be wind, turbulence, vortex or radial.");
      ForceRec r;
      for (int a = 0; a < 3; a++) {
        r.axis[a] = (float)fc.vector[a];
        r.center[a] = (float)fc.center[a];
      }
      r.strength = (float)fc.strength;
      r.kind = (float)fc.type;
      r.scale = (float)fc.scale;
      r.lift = (float)fc.lift;
      r.inward = (float)fc.inward;
      r.targeted = fc.targeted != 0;
      r.firstRegion = (int)hostRegions.size();
      r.regionCount = 0;
      r.meshTargets = 0;
      for (int k = 0; fc.targeted && k < fc.source_count; k++) {
        const fp_source &s = fc.sources[k];
        if (s.shape == FP_SHAPE_MESH) {
          r.meshTargets = 1;  // its cells come from the coverage table
          anyMeshTargets = true;
          continue;
        }
        r.regionCount++;
        Region g;
        for (int a = 0; a < 3; a++) {
          g.center[a] = (float)s.position[a];
          g.size[a] = (float)s.size[a];
        }
        g.kind = s.shape == FP_SHAPE_BOX ? 3.0f : 1.0f;
        for (int a = 0; a < 3; a++) g.invSize[a] = 1.0f / std::max(g.size[a], 0.03f);
        hostRegions.push_back(g);
      }
      hostForces.push_back(r);
    }
    if (anyMeshTargets) {
      try {
        coverageHost = mesh::forceCoverage(f, velocitySpacing(), brickCount(), sV, [&](int bx, int by, int bz) {
          int slot = dir.slotAt(bx >> 2, by >> 2, bz >> 2);
          return slot < 0 ? -1 : slot * 64 + ((bz & 3) * 4 + (by & 3)) * 4 + (bx & 3);
        });
      } catch (const std::invalid_argument &e) {
        throw InvalidArgument(e.what());
      }
      upload(forceCoverage, coverageHost);
    }
    hasForceCoverage = anyMeshTargets;
    upload(emitters, hostEmitters);
    upload(colliders, hostColliders);
    upload(forces, hostForces);
    upload(regions, hostRegions);
    emitterCount = (int)hostEmitters.size();
    forceCount = (int)hostForces.size();
    return changed;
  }

  void setParams(const fp_frame &f, double dt) {
    const fp_settings &s = f.simulation;
    P.hf = (float)voxel;
    P.hv = (float)velocitySpacing();
    P.invHf = 1.0f / P.hf;
    P.invHv = 1.0f / P.hv;
    P.coarse = smokeDiv > 1 ? 1 : 0;
    P.hs = (float)(voxel * smokeDiv);
    P.invHs = 1.0f / P.hs;
    P.smokeRatio = smokeRatio;
    P.meshRecords = meshHost.count();
    P.dt = (float)dt;
    P.buoyancy = (float)s.buoyancy;
    P.cooling = (float)s.cooling;
    P.smokeWeight = (float)s.smoke_weight;
    P.damping = (float)s.velocity_damping;
    P.dissipation = (float)s.dissipation;
    P.vorticity = (float)s.vorticity;
    P.cutoff = (float)cutoff;
    P.lifetimeZero = !mayHaveFlame;
    P.time = (float)time;
    P.seed = (float)s.seed;
    std::memcpy(&P.seedBits, &P.seed, 4);
    P.colliders = colliderCount;
    for (int a = 0; a < 3; a++) {
      P.solidLo[a] = solidLo[a];
      P.solidHi[a] = solidHi[a];
    }
    P.ground = ground;
    fp_fuel fuel = f.has_fuel ? f.fuel : fp_fuel{0, 0.5, 4.0};
    P.ignition = (float)fuel.ignition_heat;
    P.burnRate = (float)fuel.burn_rate;
    P.fuelEnabled = fuel.enabled != 0;
    P.emitterCount = emitterCount;
    P.forceCount = forceCount;
    P.maccormack = maccormack;
    P.burnKeep = fexp(-P.burnRate * P.dt);
    P.lossS = fexp(-P.dissipation * P.dt);
    P.lossH = fexp(-P.cooling * P.dt);
    P.halfS = fexp(-P.dissipation * P.dt * 0.5f);
    P.halfH = fexp(-P.cooling * P.dt * 0.5f);
    P.damp = fexp(-P.damping * P.dt);
    P.lifeStep = P.fdt / P.lifespan;
  }

  AllocArgs allocArgs() {
    AllocArgs a;
    a.dir = dirArgs();
    a.L = layout();
    a.P = P;
    a.brickCount = brickCount();
    a.blocks = (brickCount() + 255) / 256;
    a.pages = pages.p;
    a.entries = entries.p;
    a.records = records.p;
    a.brickState = brickState.p;
    a.activity = activity.p;
    a.stateTmp = stateTmp.p;
    a.rankTmp = rankTmp.p;
    a.blockSums = blockSums.p;
    a.blockPrefix = blockPrefix.p;
    a.freeSlots = freeSlots.p;
    a.poolState = poolState.p;
    a.held = std::max(0, poolSlots - slotLimit());
    a.counts = counts.p;
    a.active = active.p;
    a.zeroList = zeroList.p;
    a.emitters = emitters.p;
    a.colliders = colliders.p;
    a.solids = solids.p;
    a.colliderCount = colliderCount;
    a.buildSolids = (colliderCount > 0 || solidsStale) ? 1 : 0;
    a.solidsEvery = solidsStale ? 1 : 0;
    a.colliderMasks = colliderMasks.p;
    a.meshTables = meshHost.count() ? meshTables.p : nullptr;
    a.p = pools();
    return a;
  }
  Pools pools() {
    Pools p;
    p.F = F.p;
    p.Fp = Fp.p;
    p.Fn = Fn.p;
    p.Don = Don.p;
    p.Vtr = Vtr.p;
    p.S = S.p;
    p.Sp = Sp.p;
    p.Sn = Sn.p;
    p.E = E.p;
    p.V = V.p;
    p.Vadv = Vadv.p;
    p.Vcor = Vcor.p;
    p.Vf = Vf.p;
    p.Curl = Curl.p;
    p.D = D.p;
    p.solids = solids.p;
    p.noise = noise;
    return p;
  }
  StepArgs stepArgs() {
    StepArgs s;
    s.dir = dirArgs();
    s.L = layout();
    s.P = P;
    s.p = pools();
    s.counts = counts.p;
    s.active = active.p;
    s.activity = activity.p;
    s.brickCount = brickCount();
    s.emitters = emitters.p;
    s.forces = forces.p;
    s.regions = regions.p;
    s.meshTables = meshHost.count() ? meshTables.p : nullptr;
    s.forceCoverage = hasForceCoverage ? forceCoverage.p : nullptr;
    s.meshRecords = meshHost.count() ? reinterpret_cast<const float4 *>(meshRecs.p) : nullptr;
    return s;
  }
  // Whether each V-cycle runs its tile levels in one workgroup (k_mg_tiles_cycle).
  bool tilesFused() const { return capacity <= TC_MAX && bottom == brickLevels + 2; }
  MGArgs mgArgs() {
    MGArgs m;
    m.dir = dirArgs();
    m.L = layout();
    m.P = P;
    m.counts = counts.p;
    m.active = active.p;
    m.zeroList = zeroList.p;
    m.solids = solids.p;
    for (int L = 0; L <= bottom; L++) {
      m.lv[L].p = lv[L].p.p;
      m.lv[L].q = lv[L].q.p;
      m.lv[L].rhs = lv[L].rhs.p;
      m.lv[L].topo = lv[L].topo.p;
      m.lv[L].eflag = lv[L].eflag.p;
      m.lv[L].side = lv[L].side;
      m.lv[L].tile = lv[L].tile;
    }
    m.lv[0].p = P0.p;
    m.lv[0].q = P1.p;
    m.lv[0].rhs = D.p;
    m.bottom = bottom;
    m.tilesFused = tilesFused() ? 1 : 0;
    float inv = 1.0f / P.hv;
    float ref = 1.0f / inv;
    float nrm = ref * inv;
    m.weight = nrm * nrm;
    return m;
  }

  static void checkStep(const fp_frame &f, double dt) {
    if (!std::isfinite(dt) || dt <= 0 || dt > MAX_DT)
      throw InvalidArgument("Simulation step must be in (0, 1/30]. Use fixed substeps.");
    if (f.source_count < 0 || f.force_count < 0 || f.collider_count < 0)
      throw InvalidArgument("frame counts must be nonnegative.");
  }
  void step(const fp_frame &f, double dt) {
    checkStep(f, dt);
    substep(f, dt);
    time += dt;
    steps++;
    stepped = true;
  }

  // The host half of a step before its kernels (FluidSimulation.ts:1374-1420): tiles,
  // pool growth, the scene and the parameters.
  void prepare(const fp_frame &f, double dt) {
    if (dir.ground != ground) {
      dir.ground = ground;
      geometryDirty = true;
    }
    Clock::time_point h0 = Clock::now();
    updateTiles(f);
    hostMark("host_tiles", h0);
    h0 = Clock::now();
    int wanted = 0;
    for (int i = 0; i < f.source_count; i++)
      if (std::find(sourceIds.begin(), sourceIds.end(), f.sources[i].id) == sourceIds.end())
        wanted += sourceBricks(f.sources[i], dt);
    sourceIds.clear();
    for (int i = 0; i < f.source_count; i++) sourceIds.push_back(f.sources[i].id);
    if (wanted > 0) poolTarget = std::max(poolTarget, (int)std::ceil((usedSlots + wanted) * 1.5));
    if (poolTarget > poolSlots) growPool(poolTarget);
    if (packScene(f)) {
      solidsStale = true;
      geometryDirty = true;
    }
    packFlame(f.combustion, dt);
    fp_fuel fuel = f.has_fuel ? f.fuel : fp_fuel{0, 0.5, 4.0};
    bool anyFlame = false;
    for (int i = 0; i < f.source_count; i++) anyFlame = anyFlame || f.sources[i].flame > 0;
    mayHaveFlame = mayHaveFlame || fuel.enabled || anyFlame;
    setParams(f, dt);
    hostMark("host_scene", h0);
  }

  void substep(const fp_frame &f, double dt) {
    prepare(f, dt);
    enqueueStep();
    Clock::time_point h0 = Clock::now();
    readActivity(pinned);
    hostMark("host_activity", h0);
  }

  // Every kernel of one step, in stream order.
  void enqueueStep() {
    const Clock::time_point enqueued = Clock::now();
    const int n = brickCount();
    const unsigned blocks = (unsigned)((n + 255) / 256);
    const unsigned slotsGrid = (unsigned)(std::max(1, poolSlots) + 7) / 8 * 8;  // whole rounds over the XCDs
    AllocArgs a = allocArgs();
    int m0 = begin();
    k_build_bricks<<<blocks, 256, 0, stream>>>(a);
    k_scan_blocks<<<1, 1024, 0, stream>>>(a);
    k_free_bricks<<<blocks, 256, 0, stream>>>(a);
    k_allocate_bricks<<<blocks, 256, 0, stream>>>(a);
    if (a.buildSolids && a.solidsEvery) k_build_solids<<<slotsGrid, 64, 0, stream>>>(a);
    const int freshSolids = a.buildSolids && !a.solidsEvery ? 1 : 0;  // built with the zeroing
    solidsStale = false;
    const int zchunks = std::max(1, nF / 1024);
    const unsigned linkBlocks = (slotsGrid + LINK_BRICKS - 1) / LINK_BRICKS;
    const unsigned zeroBlocks = std::min(slotsGrid * (unsigned)zchunks, FRESH_GRID);
    k_link_fresh<<<linkBlocks + zeroBlocks, 256, 0, stream>>>(a, (int)linkBlocks, zchunks, freshSolids);
    end(m0, "allocation");

    StepArgs s = stepArgs();
    // Velocity-grid kernels loop over a brick's nV cells: lanes for every cell up to 256.
    const unsigned vLanes = (unsigned)std::min(256, std::max(64, nV));
    int m1 = begin();
    k_advect_velocity<<<slotsGrid, vLanes, 0, stream>>>(s);
    end(m1, "advect_velocity");
    int m2 = begin();
    k_correct_velocity<<<slotsGrid, vLanes, 0, stream>>>(s);
    end(m2, "correct_velocity");
    const int fchunks = std::max(1, nF / 256);
    int m3 = begin();
    k_advect_scalars<<<slotsGrid * fchunks, 256, 0, stream>>>(s, fchunks);
    end(m3, "advect_scalars");
    int m4 = begin(true);
    // Field evolution has no LDS staging to share across a workgroup: 64-lane workgroups
    // (one wavefront) wait on no one else's loads (0.298 -> 0.286 ms at the four-fires state).
    const int echunks = std::max(1, nF / 64);
    k_evolve_fields<<<slotsGrid * echunks, 64, 0, stream>>>(s, echunks);
    end(m4, "evolve_fields");
    std::swap(F, Fn);
    s.p.F = F.p;
    s.p.Fn = Fn.p;
    if (smokeDiv > 1) {
      // Coarse smoke after the fine evolution: the backtrace planes are free again.
      int ms = begin();
      k_advect_smoke<<<slotsGrid, 64, 0, stream>>>(s);
      k_evolve_smoke<<<slotsGrid, 64, 0, stream>>>(s);
      end(ms, "smoke");
      std::swap(S, Sn);
      s.p.S = S.p;
      s.p.Sn = Sn.p;
    }
    // The activity readback (FluidSimulation.ts:1514-1523): the counts, pool state and tile
    // content flags are final once the fields have evolved, so they are copied out here and
    // the host's next step overlaps the forces, the pressure solve and the projection.
    // With vorticity the curl shares the readback's launch (its workgroups first).
    int m5 = begin();
    if (P.vorticity != 0.f) {
      const unsigned rbBlocks = (unsigned)((capacity + (int)vLanes - 1) / (int)vLanes);
      k_curl_readback<<<slotsGrid + rbBlocks, vLanes, 0, stream>>>(s, (int)slotsGrid, counts.p, poolState.p,
                                                                  activity.p + n, capacity, readback.p);
    } else {
      k_readback<<<(unsigned)((capacity + 255) / 256 + 1), 256, 0, stream>>>(counts.p, poolState.p, activity.p + n,
                                                                             capacity, readback.p);
    }
    HIPCHECK(hipMemcpyAsync(pinned, readback.p, (4 + (size_t)capacity) * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (!readbackEvent) HIPCHECK(hipEventCreateWithFlags(&readbackEvent, hipEventDisableTiming));
    HIPCHECK(hipEventRecord(readbackEvent, stream));
    // The velocity predictor is dead after correction: it holds the cell forces.
    k_cell_forces<<<slotsGrid, vLanes, 0, stream>>>(s);
    k_forces<<<slotsGrid, vLanes, 0, stream>>>(s);
    k_divergence<<<slotsGrid, vLanes, 0, stream>>>(s);
    end(m5, "forces_divergence");
    int m6 = begin();
    solvePressure(slotsGrid);
    end(m6, "pressure");
    int m7 = begin();
    s.pressure = latestPressure;
    s.pressureLo = latestPressureLo;
    k_project<<<slotsGrid, vLanes, 0, stream>>>(s);
    end(m7, "project");
    exportSolverField();
    hostMark("host_enqueue", enqueued);
    Clock::time_point w0 = Clock::now();
    if (profiling && profileLevel != 2) {
      HIPCHECK(hipStreamSynchronize(stream));  // every stage's interval is read below
    } else {
      // The step returns once its readback has landed; the rest stays queued on the stream
      // (drain() before anything reads or changes device state).
      HIPCHECK(hipEventSynchronize(readbackEvent));
      tailPending = true;
    }
    hostMark("host_wait", w0);
    HIPCHECK(hipGetLastError());
    if (profiling) collect();
  }

  // After a tile-set change, the topology the change can reach. With colliders: the bricks
  // of the live tiles within a brick of the colliders' box of field cells, listed on the
  // host, then every brick of them that holds a slot tested on the device (nearSolids). With
  // tile levels: their topology at the tiles created by the last table upload and, with
  // colliders, at the live tiles within a tile of the colliders' box (a tile-level cell's
  // neighbor samples lie up to a tile away). Both lists go up in one copy. With `fresh`
  // workgroups for the zeroed list's bricks, that pass joins the same launch (128 lanes).
  void movedTileTopology(const MGArgs &m, unsigned lanes, unsigned fresh) {
    tileListsHost.clear();
    const long long side = 4LL * B;  // field cells a tile side
    if (colliderCount > 0)
      for (int t = 0; t < capacity && t < (int)dir.coords.size(); t++) {
        const std::array<int, 4> &c = dir.coords[t];
        if (!c[3]) continue;
        bool near = true;
        for (int a = 0; a < 3; a++)
          if ((long long)c[a] * side + side + B - 1 < solidLo[a] || (long long)c[a] * side - B > solidHi[a]) near = false;
        if (near) tileListsHost.push_back(t);
      }
    const int nearCount = (int)tileListsHost.size();
    if (bottom >= brickLevels) {
      tileListsHost.insert(tileListsHost.end(), created.begin(), created.end());
      if (colliderCount > 0)
        for (int t = 0; t < capacity && t < (int)dir.coords.size(); t++) {
          const std::array<int, 4> &c = dir.coords[t];
          if (!c[3] || std::find(created.begin(), created.end(), t) != created.end()) continue;
          bool near = true;
          for (int a = 0; a < 3; a++)
            if ((long long)c[a] * side + 2 * side - 1 < solidLo[a] || (long long)c[a] * side - side > solidHi[a])
              near = false;
          if (near) tileListsHost.push_back(t);
        }
    }
    const int topoCount = (int)tileListsHost.size() - nearCount;
    if (!tileListsHost.empty()) {
      if (tileLists.n < tileListsHost.size()) {
        drain();
        tileLists.alloc(std::max<size_t>(tileListsHost.size(), 128));
      }
      // Read when the stream reaches the copy; each step waits for its readback, queued
      // after it, before the lists are rebuilt.
      HIPCHECK(hipMemcpyAsync(tileLists.p, tileListsHost.data(), tileListsHost.size() * sizeof(int),
                              hipMemcpyHostToDevice, stream));
    }
    if (fresh) {
      const unsigned grid = fresh + (unsigned)nearCount * 64 + (unsigned)topoCount;
      k_moved_topology<<<grid, 128, 0, stream>>>(m, (int)fresh, tileLists.p, nearCount, topoCount);
      return;
    }
    if (nearCount) k_near_topology<<<(unsigned)nearCount * 64, lanes, 0, stream>>>(m, tileLists.p, nearCount);
    if (topoCount) k_topology_tiles_list<<<(unsigned)topoCount, 128, 0, stream>>>(m, tileLists.p + nearCount, topoCount);
  }

  // PressureSolver.solve (PressureSolver.ts:371-402): one cold-start V-cycle, or the
  // defect-correction loop around V-cycles.
  void solvePressure(unsigned slotsGrid) {
    MGArgs m = mgArgs();
    const bool tiles = tilesMoved;
    tilesMoved = false;
    m.topologyAll = geometryDirty ? 1 : 0;
    m.topologyTiles = geometryDirty || tiles ? 1 : 0;
    // The brick levels hold at most nV cells per brick; their topology, every level's cells
    // at once, takes up to 1.15 nV lanes.
    const unsigned vLanes = (unsigned)std::min(256, std::max(64, nV));
    const unsigned tLanes = nV >= 256 ? 256u : 128u;
    if (geometryDirty) {
      // Every active brick, the step's fresh ones included, and every tile level at once.
      k_topology_bricks<<<slotsGrid, tLanes, 0, stream>>>(m, 0);
      if (bottom >= brickLevels) k_topology_tiles<<<capacity, 128, 0, stream>>>(m, -1);
      geometryDirty = false;
    } else {
      const unsigned freshGrid = std::min(slotsGrid * 2, FRESH_GRID);
      if (tiles) {
        // Only the tile set changed. Tiles keep their slots, and away from the colliders a
        // cell's topology depends on the floor alone: the bricks whose cells a collider may
        // reach are built again (a new brick beside a collider can close its neighbors'
        // sides), and of the tile levels the created tiles and those near the colliders.
        // At 128 lanes the fresh bricks' pass shares the launch.
        if (tLanes != 128) k_fresh_topology<<<freshGrid, tLanes, 0, stream>>>(m);
        movedTileTopology(m, tLanes, tLanes == 128 ? freshGrid : 0u);
      } else {
        k_fresh_topology<<<freshGrid, tLanes, 0, stream>>>(m);
      }
    }
    if (pressureMode == 0) {
      // The flags of every level and level 0's first iterate in one launch.
      const unsigned groups = groupGrid(0);
      const unsigned tileBlocks = bottom >= brickLevels ? (unsigned)(capacity + 3) / 4 : 0u;
      k_mg_flags_init<<<groups + tileBlocks, 256, 0, stream>>>(m, (int)groups);
      int flips = 0;
      cycle<StoreF32>(m, 0, slotsGrid, false, flips, true);
      latestPressure = (flips & 1) ? m.lv[0].q : m.lv[0].p;  // P0 or P1
      latestPressureLo = nullptr;
      lastCycles = 1;
      lastResidual = -1.0;
      return;
    }
    k_eflags_bricks<<<slotsGrid, vLanes, 0, stream>>>(m);
    if (bottom >= brickLevels) k_eflags_tiles<<<capacity, 64, 0, stream>>>(m);
    ensureCorrectionBuffers();
    if (pressureMode == 1) correct<StoreF32>(correctionArgs32(m), slotsGrid);
    else if (pressureMode == 2) correct<StoreBF16>(correctionArgs16<uint16_t>(m), slotsGrid);
    else correct<StoreF16>(correctionArgs16<_Float16>(m), slotsGrid);
    latestPressure = dcX.p;
    latestPressureLo = dcXlo.p;
  }

  // The level arrays in 16-bit storage, with the topology and flags of `m`.
  template <class T> MGArgsT<T> halfArgs(const MGArgs &m) {
    MGArgsT<T> h;
    h.dir = m.dir;
    h.L = m.L;
    h.P = m.P;
    h.counts = m.counts;
    h.active = m.active;
    h.zeroList = m.zeroList;
    h.solids = m.solids;
    h.bottom = m.bottom;
    h.weight = m.weight;
    h.topologyAll = m.topologyAll;
    h.topologyTiles = m.topologyTiles;
    h.tilesFused = m.tilesFused;
    for (int L = 0; L <= bottom; L++) {
      h.lv[L].p = reinterpret_cast<T *>(lv16[L].p.p);
      h.lv[L].q = reinterpret_cast<T *>(lv16[L].q.p);
      h.lv[L].rhs = reinterpret_cast<T *>(lv16[L].rhs.p);
      h.lv[L].topo = m.lv[L].topo;
      h.lv[L].eflag = m.lv[L].eflag;
      h.lv[L].side = m.lv[L].side;
      h.lv[L].tile = m.lv[L].tile;
    }
    return h;
  }

  void ensureCorrectionBuffers() {
    if (capacity > BOTTOM64_TILES) throw InvalidArgument("defect correction supports up to 8192 tile slots.");
    size_t cells0 = (size_t)poolSlots * nV;
    if (dcX.n != cells0) {
      drain();  // the last step's queued tail finishes before its buffers are replaced
      dcX.alloc(cells0);
      dcXlo.alloc(cells0);
      dcR.alloc(cells0);
      dcRhs.alloc(cells0);
      dcPartials.alloc(3 * (size_t)std::max(1, poolSlots));
    }
    if (!dcNorms.p) dcNorms.alloc(8);
    if (!pinnedNorms) pinnedNorms = pinnedAlloc<double>(8);
    if (pressureMode >= 2)
      for (int L = 0; L <= bottom; L++) {
        size_t n = L == 0 ? cells0 : lv[L].p.n;
        if (lv16[L].p.n != n) {
          drain();
          lv16[L].p.alloc(n);
          lv16[L].q.alloc(n);
          lv16[L].rhs.alloc(n);
        }
      }
  }
  template <class T> void setCorrection(MGArgsT<T> &m) {
    m.dcB = D.p;
    m.dcX = dcX.p;
    m.dcXlo = dcXlo.p;
    m.dcR = dcR.p;
    m.dcPartials = dcPartials.p;
    m.dcNorms = dcNorms.p;
    m.dcTol = pressureTol;
    m.dcMaxCycles = pressureMaxCycles;
    m.dcSlots = poolSlots;
  }
  // The solve's arguments in the storage of pressure mode 1 (FP32 levels, level 0's
  // right-hand side apart from b) or 2 and 3 (16-bit levels).
  MGArgs correctionArgs32(const MGArgs &m) {
    MGArgs mf = m;
    mf.lv[0].rhs = dcRhs.p;
    setCorrection(mf);
    return mf;
  }
  template <class T> MGArgsT<T> correctionArgs16(const MGArgs &m) {
    MGArgsT<T> h = halfArgs<T>(m);
    setCorrection(h);
    return h;
  }

  // Defect correction: x = 0; r = b - A x; while |r| / |b| > tol, solve A e = r / s with
  // one V-cycle in the levels' storage (s the power of two under max |r|) and add s e to x.
  // Norms and the stopping test are FP64 on the device; the host reads the outcome to
  // decide whether to enqueue the next cycle.
  template <class St> void correct(MGArgsT<typename St::T> m, unsigned slotsGrid) {
    k_dc_start<St><<<slotsGrid, 256, 0, stream>>>(m, 0);
    lastCycles = 0;
    lastResidual = 1.0;
    for (int it = 0; it <= pressureMaxCycles; it++) {
      k_dc_residual<St><<<slotsGrid, 256, 0, stream>>>(m, 0);
      k_dc_norms<St><<<1, 256, 0, stream>>>(m, it);
      HIPCHECK(hipMemcpyAsync(pinnedNorms, dcNorms.p, 8 * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      lastResidual = pinnedNorms[4];
      if (pinnedNorms[5] != 0.0) break;
      k_dc_scale<St><<<slotsGrid, 256, 0, stream>>>(m, 0);
      int flips = 0;
      cycle<St>(m, 0, slotsGrid, true, flips);
      k_dc_update<St><<<slotsGrid, 256, 0, stream>>>(m, flips << 8);
      lastCycles = it + 1;
    }
  }

  // One V-cycle from level L (PressureSolver.ts:303-363). `flips` tracks which array holds
  // each level's iterate: a sweep writes the other array and flips its level's bit.
  // Workgroups of a grouped multigrid launch over brick level L: groups of groupSize(cells,
  // 256) bricks (the kernels' grouping) for at most poolSlots bricks, in whole XCD rounds.
  unsigned groupGrid(int L) const {
    const int sd = b >> L, n = sd * sd * sd;
    const int k = n >= 256 ? 1 : std::min(256 / n, MG_GROUP);
    const int groups = (std::max(1, poolSlots) + k - 1) / k;
    return (unsigned)((groups + 7) / 8 * 8);
  }
  template <class St> void relax(MGArgsT<typename St::T> &m, int L, unsigned slotsGrid, int &flips) {
    if (L < brickLevels) k_mg_relax_bricks<St><<<groupGrid(L), 256, 0, stream>>>(m, L | flips << 8);
    else k_mg_relax_tiles<St><<<capacity, 64, 0, stream>>>(m, L | flips << 8);
    flips ^= 1 << L;
  }
  // With `initDone`, level 0's first iterate is already in place (k_mg_flags_init).
  template <class St>
  void cycle(MGArgsT<typename St::T> &m, int L, unsigned slotsGrid, bool bottom64, int &flips, bool initDone = false) {
    if (L == bottom) {
      if (bottom64) k_mg_bottom64<St><<<1, 1024, 0, stream>>>(m, flips << 8);
      else if (capacity <= BOTTOM64_TILES) k_mg_bottom<St><<<1, 1024, 0, stream>>>(m, flips << 8);
      else
        for (int i = 0; i < BOTTOM_ITERATIONS; i++) relax<St>(m, L, slotsGrid, flips);
      return;
    }
    if (L == brickLevels) {
      // The first tile level: its sweeps and the restriction below it, the levels below (in
      // one workgroup when the tile slots fit), then the prolongation into it and its sweeps.
      // Each halo kernel leaves the level's iterate in its other array.
      k_mg_tiles_down<St><<<capacity, 256, 0, stream>>>(m, flips);
      flips ^= 1 << L;
      if (m.tilesFused) k_mg_coarse_cycle<St><<<1, TC_LANES, 0, stream>>>(m, flips | (bottom64 ? 1 << 16 : 0));
      else cycle<St>(m, L + 1, slotsGrid, bottom64, flips);
      k_mg_tiles_up<St><<<capacity, 256, 0, stream>>>(m, flips);
      flips ^= 1 << L;
      return;
    }
    if (L >= 1 && L == brickLevels - 1 && m.lv[L].side == 2) {
      // The last brick level at 2^3 cells a brick: its two sweeps and the restriction below
      // it, then after the tile levels the prolongation into it and its three sweeps, each
      // one launch of a tile of bricks and its halo per workgroup (the result in the other
      // array).
      k_mg_bricks_down<St><<<capacity, 256, 0, stream>>>(m, flips);
      flips ^= 1 << L;
      cycle<St>(m, L + 1, slotsGrid, bottom64, flips);
      k_mg_bricks_up<St><<<capacity, 256, 0, stream>>>(m, flips);
      flips ^= 1 << L;
      return;
    }
    if (L == 0) {
      if (!initDone) k_mg_init<St><<<groupGrid(0), 256, 0, stream>>>(m, flips << 8);
    } else {
      relax<St>(m, L, slotsGrid, flips);
      relax<St>(m, L, slotsGrid, flips);
    }
    if (L + 1 < brickLevels) k_mg_restrict_bricks<St><<<groupGrid(L + 1), 256, 0, stream>>>(m, (L + 1) | flips << 8);
    else k_mg_restrict_tiles<St><<<capacity, 64, 0, stream>>>(m, (L + 1) | flips << 8);
    cycle<St>(m, L + 1, slotsGrid, bottom64, flips);
    if (L < brickLevels) k_mg_prolong_bricks<St><<<groupGrid(L), 256, 0, stream>>>(m, L | flips << 8);
    else k_mg_prolong_tiles<St><<<capacity, 64, 0, stream>>>(m, L | flips << 8);
    flips ^= 1 << L;
    relax<St>(m, L, slotsGrid, flips);
    relax<St>(m, L, slotsGrid, flips);
    relax<St>(m, L, slotsGrid, flips);
  }

  // readActivity (FluidSimulation.ts:1701-1746), applied at once.
  void readActivity(const int *words) {
    int bricks = words[0], free = words[1], failures = words[2];
    lastZeroed = words[3];
    contentTiles.clear();
    for (int t = 0; t < capacity; t++) {
      uint32_t flags = (uint32_t)words[4 + t];
      if (flags && t < (int)dir.coords.size() && dir.coords[t][3])
        contentTiles.push_back({{dir.coords[t][0], dir.coords[t][1], dir.coords[t][2]}, flags});
    }
    lastFailures = failures;
    freeCount = free;
    activeBricks = bricks;
    if (steps < poolGrownAt) return;
    usedSlots = poolSlots - free;
    int limit = slotLimit();
    poolLimited = failures > 0 && poolSlots >= limit;
    if (failures > 0 && poolSlots < limit)
      poolTarget = std::max(std::max(poolTarget, (int)std::ceil((usedSlots + (double)failures) * 1.5)), poolSlots + side);
  }

  void reset() {
    allocate();
    time = 0;
    steps = 0;
    mayHaveFlame = false;
    poolGrownAt = 0;
    activeBricks = 0;
    stepped = false;
    rc.lastKey.clear();  // SceneVolumeRenderer.reset -> invalidate (SceneVolumeRenderer.ts:254-261)
  }

  // ---- render (VolumeLighting.encode, VolumeLighting.ts:249-375; the camera march) ----
  RenderParams renderParams(const fp_camera &cam, const fp_render_settings &s) const {
    auto finite = [](double v, const char *name) {
      if (!std::isfinite(v)) throw InvalidArgument(std::string(name) + " must be finite.");
      return (float)v;
    };
    if (cam.width < 1 || cam.height < 1 || cam.width > 16384 || cam.height > 16384)
      throw InvalidArgument("The image must be 1 to 16384 pixels a side.");
    int ld = s.lighting_divisor;
    if (ld != 1 && ld != 2 && ld != 4) throw InvalidArgument("rendering.lightingDivisor must be 1, 2 or 4.");
    if (B / ld < 1) throw InvalidArgument("rendering.lightingDivisor exceeds the brick size.");
    if (s.filter < 0 || s.filter > 2) throw InvalidArgument("rendering.filter must be trilinear, quadratic, cubic.");
    if (s.ray_steps < 1) throw InvalidArgument("rendering.raySteps must be a positive integer.");
    RenderParams P{};
    P.voxel = (float)voxel;
    P.shift = sF;
    P.lightDivisor = ld;
    P.lightShift = sF - (ld == 4 ? 2 : ld == 2 ? 1 : 0);
    P.brickSize = P.voxel * (float)B;
    P.texel = P.voxel * (float)ld;
    P.flameOpacity = finite(s.flame_opacity, "flame.opacity");
    P.optical[0] = finite(s.smoke_density, "smoke.density");
    P.optical[1] = finite(s.fire_intensity, "flame.brightness");
    P.optical[2] = finite(s.shadow_density, "smoke.shadowDensity");
    P.optical[3] = finite(s.scattering, "smoke.scattering");
    for (int a = 0; a < 3; a++) {
      P.light[a] = finite(s.sun_direction[a], "light direction");
      P.smoke[a] = finite(s.smoke_color[a], "smoke.color");
      P.fire[a] = finite(s.fire_color[a], "flame.color");
      P.sceneLight[a] = finite(s.sun_color[a], "light color");
      P.ambient[a] = finite(s.ambient[a], "ambient light");
    }
    P.light[3] = finite(s.sun_intensity, "light intensity");
    P.sootGlow = finite(s.soot_glow, "flame.sootGlow");
    P.fire[3] = finite(s.temperature, "flame.temperature");
    P.filter = (float)s.filter;
    P.raySteps = (float)s.ray_steps;
    P.time = (float)time;
    for (int i = 0; i < 16; i++) P.ivp[i] = finite(cam.inverse_view_projection[i], "inverse view-projection");
    P.width = cam.width;
    P.height = cam.height;
    return P;
  }

  RenderArgs renderArgs(const RenderParams &P) const {
    RenderArgs r{};
    r.B = B;
    r.sF = sF;
    r.side = side;
    r.nF = nF;
    r.E = B + 1;
    r.nR = r.E * r.E * r.E;
    r.lside = 1 << P.lightShift;
    r.LE = r.lside + 1;
    r.nL = r.LE * r.LE * r.LE;
    r.perSide = B / 8;
    r.nO = r.perSide * r.perSide * r.perSide;
    r.boxes = rc.boxes.p;
    r.boxCount = rc.boxCount;
    r.table = rc.table.p;
    r.tiles = rc.tiles.p;
    r.pages = pages.p;
    r.counts = counts.p;
    r.active = active.p;
    r.zeroList = zeroList.p;
    r.brickCount = brickCount();
    r.F = F.p;
    r.R = rc.R.p;
    r.Lt = rc.L.p;
    r.O = rc.O.p;
    r.lut = rc.lut.p;
    r.P = P;
    r.coarse = smokeDiv > 1 ? 1 : 0;
    r.sS = smokeShift;
    r.bs = smokeEdge;
    r.nS = nS;
    r.ES = smokeEdge + 1;
    r.nRS = r.ES * r.ES * r.ES;
    r.S = S.p;
    r.RS = rc.RS.p;
    r.dir = dirArgs();
    return r;
  }

  static std::string settingsKey(const RenderParams &p) {
    char buf[512];
    snprintf(buf, sizeof buf, "%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%a,%d", p.optical[0], p.optical[2],
             p.optical[1], p.flameOpacity, p.fire[3], p.fire[0], p.fire[1], p.fire[2], p.smoke[0], p.smoke[1],
             p.smoke[2], p.sootGlow, p.light[0], p.light[1], p.light[2], p.light[3] > 0.0f ? 1 : 0);
    return buf;
  }

  // VolumeLighting.sync + sizePools + encode: bring the page tables, render field,
  // occupancy and light up to the latest step.
  void renderPrepare(const RenderParams &P) {
    if (!rc.lut.p) {
      std::vector<float> t = render::blackbody();
      rc.lut.alloc(t.size() / 4);
      HIPCHECK(hipMemcpyAsync(rc.lut.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    }
    if (rboxes.version != rc.version) {
      rc.version = rboxes.version;
      rc.boxCount = (int)rboxes.boxes.size();
      rc.boxWords.assign((size_t)std::max(1, rc.boxCount) * 8, 0);
      int entries = 0;
      for (int i = 0; i < rc.boxCount; i++) {
        const auto &b = rboxes.boxes[i];
        int32_t *w = &rc.boxWords[(size_t)i * 8];
        int size[3];
        for (int a = 0; a < 3; a++) size[a] = b.hi[a] - b.lo[a] + 1;
        w[0] = b.lo[0], w[1] = b.lo[1], w[2] = b.lo[2], w[3] = entries;
        w[4] = size[0], w[5] = size[1], w[6] = size[2], w[7] = 0;
        entries += size[0] * size[1] * size[2];
      }
      rc.pageCount = entries;
      if (rc.boxes.n < rc.boxWords.size() / 4) rc.boxes.alloc(rc.boxWords.size() / 4);
      HIPCHECK(hipMemcpyAsync(rc.boxes.p, rc.boxWords.data(), rc.boxWords.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, stream));
      if (rc.tiles.n < rboxes.tiles.size() / 4) rc.tiles.alloc(rboxes.tiles.size() / 4);
      HIPCHECK(hipMemcpyAsync(rc.tiles.p, rboxes.tiles.data(), rboxes.tiles.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, stream));
      size_t tableWords = (size_t)std::max(1, entries);
      if (rc.table.n < tableWords) rc.table.alloc(tableWords);
      k_fill_int<<<(unsigned)((tableWords + 255) / 256), 256, 0, stream>>>(rc.table.p, -1, tableWords);
      HIPCHECK(hipStreamSynchronize(stream));  // rc.boxWords and rboxes.tiles are host memory
      rc.stale = true;
    }
    bool resized = poolSlots != rc.poolSlots || P.lightDivisor != rc.lightDivisor;
    if (resized) {
      rc.poolSlots = poolSlots;
      rc.lightDivisor = P.lightDivisor;
      RenderArgs r = renderArgs(P);
      rc.R.alloc((size_t)poolSlots * r.nR);
      rc.L.alloc((size_t)poolSlots * r.nL);
      rc.O.alloc((size_t)poolSlots * r.nO);
      if (smokeDiv > 1) {
        rc.RS.alloc((size_t)poolSlots * r.nRS);
        HIPCHECK(hipMemsetAsync(rc.RS.p, 0, rc.RS.bytes(), stream));
      }
      HIPCHECK(hipMemsetAsync(rc.R.p, 0, rc.R.bytes(), stream));
      HIPCHECK(hipMemsetAsync(rc.L.p, 0, rc.L.bytes(), stream));
      HIPCHECK(hipMemsetAsync(rc.O.p, 0, rc.O.bytes(), stream));
    }
    bool changed = rc.stale;
    std::string key = settingsKey(P);
    bool stepNew = steps != rc.lastStep;
    bool rebuild = changed || resized;
    bool smooth = stepNew || rebuild;
    bool pagesDue = stepNew || changed;
    bool occupancy = rc.occupancyStep != steps || rebuild;
    bool lightDue = steps < rc.lastLightStep || steps - rc.lastLightStep >= 2;
    bool buildLight = (stepNew && lightDue) || rebuild || key != rc.lastKey;
    bool fresh = stepNew && !buildLight;
    if (rc.boxCount == 0 || !(smooth || pagesDue || occupancy || buildLight || fresh)) {
      rc.lastWork = -1;
      return;
    }
    RenderArgs r = renderArgs(P);
    unsigned slotsGrid = (unsigned)(std::max(1, poolSlots) + 7) / 8 * 8;  // whole rounds over the XCDs
    int t = begin();
    if (pagesDue) k_export_pages<<<(unsigned)((brickCount() + 255) / 256), 256, 0, stream>>>(r);
    if (smooth) {
      int chunks = (r.nR + 1023) / 1024;
      bool staged = B <= 16 && smokeDiv == 1;
      size_t lds = staged ? (size_t)(B + 3) * (B + 3) * (B + 3) * sizeof(float) : 0;
      k_smooth_smoke<<<slotsGrid * chunks, 256, lds, stream>>>(r, chunks, staged ? 1 : 0);
      if (smokeDiv > 1) k_smooth_coarse<<<slotsGrid, 256, 0, stream>>>(r);
    }
    if (occupancy) k_occupancy<<<slotsGrid, 256, 0, stream>>>(r);
    end(t, "render_prepare");
    t = begin();
    if (buildLight) k_light<<<slotsGrid, 64, 0, stream>>>(r, 0);
    else if (fresh) k_light<<<slotsGrid, 64, 0, stream>>>(r, 1);
    end(t, "render_light");
    rc.lastWork = buildLight ? 1 : fresh ? 2 : 0;
    if (occupancy) rc.occupancyStep = steps;
    if (buildLight) {
      rc.lastLightStep = steps;
      rc.lastKey = key;
    }
    rc.lastStep = steps;
    rc.stale = false;
  }

  void renderImage(const fp_camera &cam, const fp_render_settings &s, float *out, bool deviceOut) {
    RenderParams P = renderParams(cam, s);
    renderPrepare(P);
    size_t pixels = (size_t)P.width * P.height;
    float4 *target = reinterpret_cast<float4 *>(out);
    if (!deviceOut) {
      if (rc.image.n < pixels) rc.image.alloc(pixels);
      target = rc.image.p;
    }
    RenderArgs r = renderArgs(P);
    dim3 grid((unsigned)((P.width + 15) / 16), (unsigned)((P.height + 15) / 16));
    if (!rc.sampleCounts.p) rc.sampleCounts.alloc(4);
    k_fill_u32<<<1, 256, 0, stream>>>(reinterpret_cast<uint32_t *>(rc.sampleCounts.p), 0u, 8);
    const float4 *slices = nullptr;
    int sw = 0, sh = 0;
    if (s.half_resolution) {
      // VolumeSlices.march (VolumeSlices.ts:83-131): half the size on each axis, 8 depths.
      sw = std::max(1, (P.width + 1) / 2);
      sh = std::max(1, (P.height + 1) / 2);
      size_t n = (size_t)sw * sh * 8;
      if (rc.slices.n < n) rc.slices.alloc(n);
      const int tiles = ((sw + 15) / 16) * ((sh + 15) / 16);
      if (rc.tileCost.n < (size_t)tiles) {
        rc.tileCost.alloc(tiles);
        rc.tileOrder.alloc(tiles);
        rc.orderTiles = 0;
      }
      const int *order = rc.orderTiles == tiles ? rc.tileOrder.p : nullptr;
      int t0 = begin();
      k_slices<<<(unsigned)tiles, 256, 0, stream>>>(r, rc.slices.p, sw, sh, order, rc.tileCost.p, rc.sampleCounts.p);
      end(t0, "render_slices");
      k_order_tiles<<<1, 1024, 0, stream>>>(rc.tileCost.p, tiles, rc.tileOrder.p);
      rc.orderTiles = tiles;
      slices = rc.slices.p;
    }
    int t = begin();
    k_march<<<grid, 256, 0, stream>>>(r, target, slices, sw, sh, rc.sampleCounts.p);
    end(t, "render_march");
    HIPCHECK(hipGetLastError());
    if (!deviceOut) {
      HIPCHECK(hipMemcpyAsync(out, rc.image.p, pixels * sizeof(float4), hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
    } else {
      // The render kernels stay queued and still read the pools; the next entry point that
      // reads or replaces device state waits for them.
      tailPending = true;
    }
    if (profiling) {
      HIPCHECK(hipStreamSynchronize(stream));
      collect();
    }
  }

  // SceneLights.sample (SceneLights.ts:127-218) over the latest step's fields: per anchor,
  // the luminance of the flames nearest it, their luminance-weighted offset and their rgb.
  void sceneLights(const fp_render_settings &s, const double *xyz, int count, float *out) {
    if (count < 0 || count > 65536) throw InvalidArgument("anchor count must be 0 to 65536.");
    for (int i = 0; i < 3 * count; i++)
      if (!std::isfinite(xyz[i])) throw InvalidArgument("anchor positions must be finite.");
    fp_camera cam{};
    cam.width = cam.height = 1;
    RenderParams P = renderParams(cam, s);
    std::fill(out, out + (size_t)8 * count, 0.0f);
    const int n = (int)std::min<int64_t>(activeBricks, (int64_t)brickCount());
    if (count == 0 || rboxes.boxes.empty() || n == 0 || !stepped) return;
    if (!rc.lut.p) {
      std::vector<float> t = render::blackbody();
      rc.lut.alloc(t.size() / 4);
      HIPCHECK(hipMemcpyAsync(rc.lut.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    }
    std::vector<float4> a(count);
    for (int i = 0; i < count; i++) a[i] = make_float4((float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2], 0.f);
    const int blocks = (n + LIGHT_BLOCK - 1) / LIGHT_BLOCK;
    if (rc.anchors.n < (size_t)count) rc.anchors.alloc(count);
    if (rc.partials.n < (size_t)n * 8) rc.partials.alloc((size_t)n * 8);
    if (rc.brickAnchor.n < (size_t)n) rc.brickAnchor.alloc(n);
    if (rc.blockSums.n < (size_t)blocks * count * 8) rc.blockSums.alloc((size_t)blocks * count * 8);
    if (rc.sums.n < (size_t)count * 8) rc.sums.alloc((size_t)count * 8);
    HIPCHECK(hipMemcpyAsync(rc.anchors.p, a.data(), count * sizeof(float4), hipMemcpyHostToDevice, stream));
    RenderArgs r = renderArgs(P);
    int t = begin();
    k_light_bricks<<<(unsigned)n, 64, 0, stream>>>(r, tileRec.p, rc.anchors.p, count, rc.partials.p, rc.brickAnchor.p);
    k_light_blocks<<<(unsigned)blocks, LIGHT_BLOCK, 0, stream>>>(counts.p, rc.partials.p, rc.brickAnchor.p, count,
                                                                 rc.blockSums.p);
    k_light_total<<<1, LIGHT_BLOCK, 0, stream>>>(counts.p, rc.blockSums.p, count, rc.sums.p);
    end(t, "scene_lights");
    HIPCHECK(hipMemcpyAsync(out, rc.sums.p, (size_t)count * 8 * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipGetLastError());
    if (profiling) collect();
  }

  // sceneLights without allocating: the same kernels on buffers in `scratch` (the
  // device address `base`), for at most NOALLOC_MAX_ANCHORS anchors over at most
  // NOALLOC_MAX_BRICKS active bricks. Every scratch word a call reads it writes first, so
  // nothing needs clearing between calls; calls that overlap on one GPU share the array
  // and race.
  void sceneLightsNoAlloc(const fp_render_settings &s, const double *xyz, int count, float *out,
                          unsigned char *base) {
    if (count < 0 || count > NOALLOC_MAX_ANCHORS)
      throw InvalidArgument("the allocation-free scene lights take 0 to 64 anchors.");
    for (int i = 0; i < 3 * count; i++)
      if (!std::isfinite(xyz[i])) throw InvalidArgument("anchor positions must be finite.");
    fp_camera cam{};
    cam.width = cam.height = 1;
    RenderParams P = renderParams(cam, s);
    std::fill(out, out + (size_t)8 * count, 0.0f);
    const int n = (int)std::min<int64_t>(activeBricks, (int64_t)brickCount());
    if (n > NOALLOC_MAX_BRICKS)
      throw InvalidArgument("the allocation-free scene lights take at most 65536 active bricks.");
    if (count == 0 || rboxes.boxes.empty() || n == 0 || !stepped) return;
    float4 a[NOALLOC_MAX_ANCHORS];
    for (int i = 0; i < count; i++) a[i] = make_float4((float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2], 0.f);
    float4 *anchors = reinterpret_cast<float4 *>(base + NA_ANCHORS);
    float *partials = reinterpret_cast<float *>(base + NA_PARTIALS);
    int *brickAnchor = reinterpret_cast<int *>(base + NA_BRICK_ANCHOR);
    float *blockSums = reinterpret_cast<float *>(base + NA_BLOCK_SUMS);
    float *sums = reinterpret_cast<float *>(base + NA_SUMS);
    const int blocks = (n + LIGHT_BLOCK - 1) / LIGHT_BLOCK;
    HIPCHECK(hipMemcpyAsync(anchors, a, count * sizeof(float4), hipMemcpyHostToDevice, stream));
    RenderArgs r = renderArgs(P);
    r.lut = reinterpret_cast<const float4 *>(base + NA_LUT);
    k_light_bricks<<<(unsigned)n, 64, 0, stream>>>(r, tileRec.p, anchors, count, partials, brickAnchor);
    k_light_blocks<<<(unsigned)blocks, LIGHT_BLOCK, 0, stream>>>(counts.p, partials, brickAnchor, count, blockSums);
    k_light_total<<<1, LIGHT_BLOCK, 0, stream>>>(counts.p, blockSums, count, sums);
    HIPCHECK(hipMemcpyAsync(out, sums, (size_t)count * 8 * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipGetLastError());
  }

  void readRender(fp_render_tables *t) {
    HIPCHECK(hipStreamSynchronize(stream));
    t->light_work = rc.lastWork;
    t->box_count = rc.boxCount;
    t->page_count = rc.pageCount;
    if (t->boxes && rc.boxCount)
      std::memcpy(t->boxes, rc.boxWords.data(), (size_t)std::min(rc.boxCount, t->box_capacity) * 8 * sizeof(int32_t));
    if (t->page_table && rc.pageCount)
      HIPCHECK(hipMemcpy(t->page_table, rc.table.p, (size_t)std::min(rc.pageCount, t->page_capacity) * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
    if (t->blackbody && rc.lut.p) HIPCHECK(hipMemcpy(t->blackbody, rc.lut.p, rc.lut.bytes(), hipMemcpyDeviceToHost));
    // Listed bricks: active, slotted, inside their tile's box.
    int n = (int)std::min<int64_t>(activeBricks, (int64_t)brickCount());
    std::vector<int32_t> ids(n), pg(brickCount());
    if (n) HIPCHECK(hipMemcpy(ids.data(), active.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(pg.data(), pages.p, pg.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::sort(ids.begin(), ids.end());
    int E = B + 1, LE = (rc.lightDivisor > 0 ? (B / rc.lightDivisor) : B) + 1, nO = (B / 8) * (B / 8) * (B / 8);
    size_t nR = (size_t)E * E * E, nL = (size_t)LE * LE * LE;
    int listed = 0;
    for (int id : ids) {
      int page = pg[id];
      const int32_t *tile = &rboxes.tiles[(size_t)(id >> 6) * 4];
      if (page < 0 || (size_t)(id >> 6) * 4 >= rboxes.tiles.size() || tile[3] < 0 || rc.version != rboxes.version)
        continue;
      const int32_t *w = &rc.boxWords[(size_t)tile[3] * 8];
      int l = id & 63, loc[3] = {l & 3, (l >> 2) & 3, l >> 4};
      bool inside = true;
      for (int a = 0; a < 3; a++) {
        int local = tile[a] * 4 + loc[a] - w[a];
        if (local < 0 || local >= w[4 + a]) inside = false;
      }
      if (!inside) continue;
      if (listed < t->brick_capacity && rc.R.p) {
        int slot = pageToSlotHost(page);
        if (t->bricks) t->bricks[listed] = id;
        if (t->field)
          HIPCHECK(hipMemcpy(t->field + (size_t)listed * nR * 4, rc.R.p + (size_t)slot * nR, nR * sizeof(float4),
                             hipMemcpyDeviceToHost));
        if (t->light)
          HIPCHECK(hipMemcpy(t->light + (size_t)listed * nL * 4, rc.L.p + (size_t)slot * nL, nL * sizeof(float4),
                             hipMemcpyDeviceToHost));
        if (t->occupancy)
          HIPCHECK(hipMemcpy(t->occupancy + (size_t)listed * nO, rc.O.p + (size_t)slot * nO, (size_t)nO,
                             hipMemcpyDeviceToHost));
        if (t->smoke && smokeDiv > 1 && rc.RS.p) {
          const size_t nRS = (size_t)(smokeEdge + 1) * (smokeEdge + 1) * (smokeEdge + 1);
          HIPCHECK(hipMemcpy(t->smoke + (size_t)listed * nRS, rc.RS.p + (size_t)slot * nRS, nRS * sizeof(float),
                             hipMemcpyDeviceToHost));
        }
      }
      listed++;
    }
    t->brick_count = listed;
  }
  int pageToSlotHost(int page) const {
    uint32_t p = (uint32_t)page;
    return (int)(((p >> 16) * (uint32_t)side + ((p >> 8) & 255u)) * (uint32_t)side + (p & 255u));
  }

  size_t memoryBytes() const {
    size_t total = tileRec.bytes() + tileLists.bytes() + pages.bytes() + entries.bytes() + records.bytes() +
                   active.bytes() + zeroList.bytes() + brickState.bytes() + activity.bytes() + stateTmp.bytes() +
                   rankTmp.bytes() + blockSums.bytes() + blockPrefix.bytes() + freeSlots.bytes() + F.bytes() +
                   Fp.bytes() + Fn.bytes() + Don.bytes() + Vtr.bytes() + Curl.bytes() + E.bytes() + V.bytes() + Vadv.bytes() +
                   Vcor.bytes() + Vf.bytes() + D.bytes() + P0.bytes() + P1.bytes() + solids.bytes() + S.bytes() +
                   Sp.bytes() + Sn.bytes() + emitters.bytes() + colliders.bytes() + colliderMasks.bytes() +
                   solverView.bytes() + dcRhs.bytes() + dcX.bytes() + dcXlo.bytes() + dcR.bytes() +
                   dcPartials.bytes() + dcNorms.bytes();
    for (int L = 0; L <= bottom; L++)
      total += lv[L].p.bytes() + lv[L].q.bytes() + lv[L].rhs.bytes() + lv[L].topo.bytes() + lv[L].eflag.bytes() +
               lv16[L].p.bytes() + lv16[L].q.bytes() + lv16[L].rhs.bytes();
    return total;
  }
  // meshMemoryBytes and forceMemoryBytes (FluidSimulation.ts:846-851): the scene's mesh
  // emission and force data, which memoryBytes leaves out.
  size_t meshMemoryBytes() const { return meshRecs.bytes() + meshTables.bytes(); }
  size_t forceMemoryBytes() const { return forces.bytes() + regions.bytes() + forceCoverage.bytes(); }

  // encodeSolverField (FluidSimulation.ts:1550-1559): the shown field of the latest step
  // into the solver-field view, on the stream after the step.
  void exportSolverField() {
    if (!solverField || !latestPressure) return;
    const size_t n = (size_t)poolSlots * nV;
    if (solverView.n != n) {
      drain();  // the previous copy may still be queued
      solverView.alloc(n);
    }
    const float *src = solverField == FP_SOLVER_FIELD_PRESSURE ? latestPressure : D.p;
    HIPCHECK(hipMemcpyAsync(solverView.p, src, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
    tailPending = true;
  }

  // Finishes whatever the last entry point left queued (a step's tail, a device render).
  void drain() {
    if (!tailPending) return;
    tailPending = false;
    HIPCHECK(hipStreamSynchronize(stream));
  }

  void destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    tailPending = false;
    if (readbackEvent) (void)hipEventDestroy(readbackEvent);
    readbackEvent = nullptr;
    DBuf<int> *ints[] = {&tileRec, &tileLists, &pages, &entries, &records, &active, &zeroList, &poolState, &counts, &readback};
    for (auto *x : ints) x->release();
    DBuf<uint32_t> *u32s[] = {&brickState, &activity, &stateTmp, &rankTmp, &blockSums, &blockPrefix, &freeSlots, &solids};
    for (auto *x : u32s) x->release();
    DBuf<float4> *f4s[] = {&F, &Fp, &Fn, &Don, &Vtr, &Curl};
    for (auto *x : f4s) x->release();
    DBuf<float> *fs[] = {&E, &V, &Vadv, &Vcor, &Vf, &D, &P0, &P1, &S, &Sp, &Sn};
    for (auto *x : fs) x->release();
    for (int L = 0; L < MAX_LEVELS; L++) {
      lv[L].p.release();
      lv[L].q.release();
      lv[L].rhs.release();
      lv[L].topo.release();
      lv[L].eflag.release();
    }
    emitters.release();
    colliders.release();
    colliderMasks.release();
    forceCoverage.release();
    forces.release();
    regions.release();
    for (int L = 0; L < MAX_LEVELS; L++) {
      lv16[L].p.release();
      lv16[L].q.release();
      lv16[L].rhs.release();
    }
    solverView.release();
    dcRhs.release();
    dcX.release();
    dcXlo.release();
    dcR.release();
    dcPartials.release();
    dcNorms.release();
    if (pinnedNorms) pinnedRelease(pinnedNorms, 8);
    rc.boxes.release();
    rc.tiles.release();
    rc.table.release();
    rc.R.release();
    rc.RS.release();
    rc.slices.release();
    rc.tileCost.release();
    rc.sampleCounts.release();
    rc.tileOrder.release();
    rc.orderTiles = 0;
    rc.L.release();
    rc.lut.release();
    rc.image.release();
    rc.O.release();
    rc.anchors.release();
    rc.partials.release();
    rc.blockSums.release();
    rc.sums.release();
    rc.brickAnchor.release();
    for (hipEvent_t e : eventPool) (void)hipEventDestroy(e);
    eventPool.clear();
    if (pinned) pinnedRelease(pinned, pinnedWords);
    pinnedWords = 0;
    streams().give(stream);
    stream = nullptr;
  }
};

namespace {
template <class Fn> int guard(Fn fn) {
  try {
    fn();
    return FP_OK;
  } catch (const InvalidArgument &e) {
    g_error = e.what();
    return FP_EINVAL;
  } catch (const DeviceFailure &e) {
    g_error = e.what();
    return FP_EDEVICE;
  } catch (const std::bad_alloc &e) {
    g_error = "host allocation failed";
    return FP_ENOMEM;
  } catch (const std::exception &e) {
    g_error = e.what();
    return FP_ESTATE;
  }
}
}  // namespace


// A batch advances many simulations of one brick layout together (the serving form): their
// host halves run one after another, then every stage is one launch over all of them
// (k_*_batch), with one upload of the arguments and one readback. Results equal those of
// stepping each simulation alone, bit for bit: the same kernel bodies run on the same data.
struct fp_batch {
  hipStream_t stream = nullptr;
  DBuf<char> devArgs;
  char *hostArgs = nullptr;
  DBuf<int> readback;
  int *pinnedReadback = nullptr;
  size_t pinnedWords = 0;
  DBuf<double> reports;  // per simulation: defect correction's relative residual and cycles
  double *pinnedReports = nullptr;
  int *pinnedLive = nullptr;  // members whose defect correction is still running
  int grid = 2048;    // persistent workgroups per launch, a multiple of 8
  bool graphs = true;  // replay captured launch sequences
  double lastMs = 0;
  // Wall ms of the latest step's phases: host halves, argument packing, launches, the wait
  // for the device, the activity readbacks.
  double phaseMs[5] = {0, 0, 0, 0, 0};

  // The uploaded arguments: per-simulation records and block prefixes at offsets fixed by
  // the member capacity `cap`. A step with n <= cap members pads every array to cap
  // entries (records zeroed, prefixes repeating their total), so kernel parameters and
  // grids depend on the capacity alone and a captured launch sequence serves any members.
  struct Layout {
    int cap = 0;
    size_t alloc, pre, post, rb, mg, solve, blocks, one, link, slots, slots2, zero, field, tiles, levels, bytes;
  } lay;
  // The solve's virtual-block prefix of level L (MAX_LEVELS: one block per unfinished
  // member), written on the device by k_level_prefix_batch.
  size_t levelPrefix(int L) const { return lay.levels + (size_t)L * (lay.cap + 1) * sizeof(int); }
  static Layout layoutFor(int cap) {
    Layout L;
    L.cap = cap;
    size_t o = 0;
    auto take = [&](size_t bytes) {
      size_t off = (o + 255) / 256 * 256;
      o = off + bytes;
      return off;
    };
    const size_t solveBytes = std::max({sizeof(MGArgs), sizeof(MGArgsT<uint16_t>), sizeof(MGArgsT<_Float16>)});
    L.alloc = take(cap * sizeof(AllocArgs));
    L.pre = take(cap * sizeof(StepArgs));
    L.post = take(cap * sizeof(StepArgs));
    L.rb = take(cap * sizeof(ReadbackArgs));
    L.mg = take(cap * sizeof(MGArgs));
    L.solve = take(cap * solveBytes);
    size_t *prefixes[] = {&L.blocks, &L.one, &L.link, &L.slots, &L.slots2, &L.zero, &L.field, &L.tiles};
    for (size_t *q : prefixes) *q = take((cap + 1) * sizeof(int));
    L.levels = take((size_t)(MAX_LEVELS + 1) * (cap + 1) * sizeof(int));
    L.bytes = (o + 255) / 256 * 256;
    return L;
  }

  // The launch sequence of a step depends on these; one captured graph per distinct key.
  // Lanes of the velocity-grid and brick-level launches: one per velocity cell of a brick,
  // 64 to 256 (members share B and d). Any count is correct; graphs keep the one captured.
  unsigned vLanes = 256;
  using GraphKey = std::array<int, 10>;
  std::map<GraphKey, hipGraphExec_t> execs;

  void create() {
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipHostMalloc((void **)&pinnedLive, sizeof(int), hipHostMallocDefault));
    graphs = !underProfiler();
  }
  // rocprofv3 loads its tool library into the process through ROCP_TOOL_LIBRARIES (and
  // LD_PRELOAD). Its kernel tracing of replayed graphs faulted on the host side once, inside
  // the profiler's own runtime, so a batch starts with graphs off when the tool is present:
  // the kernels then run one launch at a time, in the same order with the same arguments.
  static bool underProfiler() {
    const char *tools = std::getenv("ROCP_TOOL_LIBRARIES");
    const char *preload = std::getenv("LD_PRELOAD");
    return (tools && *tools) || (preload && std::strstr(preload, "rocprofiler"));
  }
  void dropGraphs() {
    for (auto &kv : execs) (void)hipGraphExecDestroy(kv.second);
    execs.clear();
  }
  void destroy() {
    if (stream) (void)hipStreamSynchronize(stream);
    dropGraphs();
    devArgs.release();
    readback.release();
    reports.release();
    if (hostArgs) (void)hipHostFree(hostArgs);
    if (pinnedReadback) (void)hipHostFree(pinnedReadback);
    if (pinnedReports) (void)hipHostFree(pinnedReports);
    pinnedReports = nullptr;
    if (pinnedLive) (void)hipHostFree(pinnedLive);
    pinnedLive = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    hostArgs = nullptr;
    pinnedReadback = nullptr;
    stream = nullptr;
  }

  // Room for `n` members: a new layout (and new graphs) when the capacity is exceeded.
  void reserve(int n) {
    if (n <= lay.cap) return;
    HIPCHECK(hipStreamSynchronize(stream));
    dropGraphs();
    lay = layoutFor(std::max({n, 2 * lay.cap, 16}));
    if (hostArgs) (void)hipHostFree(hostArgs);
    hostArgs = nullptr;
    HIPCHECK(hipHostMalloc((void **)&hostArgs, lay.bytes, hipHostMallocDefault));
    devArgs.alloc(lay.bytes);
    reports.alloc(2 * (size_t)lay.cap);
    if (pinnedReports) (void)hipHostFree(pinnedReports);
    pinnedReports = nullptr;
    HIPCHECK(hipHostMalloc((void **)&pinnedReports, reports.bytes(), hipHostMallocDefault));
  }
  // Writes n records at `off`, zeroing the rest of the capacity.
  template <class T> void put(size_t off, const std::vector<T> &v) {
    if (!v.empty()) std::memcpy(hostArgs + off, v.data(), v.size() * sizeof(T));
    std::memset(hostArgs + off + v.size() * sizeof(T), 0, (lay.cap - v.size()) * sizeof(T));
  }
  // Prefix of per-simulation block counts, each rounded up to `round`, padded to the
  // capacity; returns the total.
  int putPrefix(size_t off, const std::vector<int> &counts, int round) {
    int *out = reinterpret_cast<int *>(hostArgs + off);
    out[0] = 0;
    for (size_t i = 0; i < counts.size(); i++) out[i + 1] = out[i] + (counts[i] + round - 1) / round * round;
    for (int i = (int)counts.size() + 1; i <= lay.cap; i++) out[i] = out[counts.size()];
    return out[counts.size()];
  }

  // Launches kernel(args, starts, cap, extra...) on `g` persistent workgroups.
  template <class Args, class Kernel, class... Extra>
  void launch(Kernel kernel, unsigned threads, size_t argOff, size_t startOff, unsigned g, Extra... extra) {
    const Args *args = reinterpret_cast<const Args *>(devArgs.p + argOff);
    const int *starts = reinterpret_cast<const int *>(devArgs.p + startOff);
    kernel<<<g, threads, 0, stream>>>(args, starts, lay.cap, extra...);
  }
  unsigned wide() const { return (unsigned)grid; }
  unsigned perSim() const { return (unsigned)std::min(grid, (lay.cap + 7) / 8 * 8); }

  void step(fp_sim *const *sims, const fp_frame *frames, int n, double dt) {
    if (n <= 0) return;
    vLanes = (unsigned)std::min(256, std::max(64, sims[0] ? sims[0]->nV : 256));
    for (int i = 0; i < n; i++) {
      if (!sims[i]) throw InvalidArgument("batch member is null.");
      if (sims[i]->B != sims[0]->B || sims[i]->d != sims[0]->d)
        throw InvalidArgument("a batch's simulations must share brick size and velocity divisor.");
      if (sims[i]->smokeDiv != 1) throw InvalidArgument("batched steps need full-resolution smoke (smokeDivisor 1).");
      if (sims[i]->pressureMode != sims[0]->pressureMode)
        throw InvalidArgument("a batch's simulations must share the pressure policy.");
      for (int j = 0; j < i; j++)
        if (sims[j] == sims[i]) throw InvalidArgument("a simulation appears twice in a batch.");
      fp_sim::checkStep(frames[i], dt);
    }
    auto t0 = std::chrono::steady_clock::now();
    // Host halves, enqueued on the batch's stream.
    std::vector<hipStream_t> own(n);
    for (int i = 0; i < n; i++) {
      fp_sim *s = sims[i];
      HIPCHECK(hipStreamSynchronize(s->stream));
      own[i] = s->stream;
      s->stream = stream;
    }
    try {
      auto p0 = std::chrono::steady_clock::now();
      for (int i = 0; i < n; i++) sims[i]->prepare(frames[i], dt);
      phaseMs[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - p0).count();
      enqueue(sims, n);
    } catch (...) {
      (void)hipStreamSynchronize(stream);
      for (int i = 0; i < n; i++) sims[i]->stream = own[i];
      throw;
    }
    for (int i = 0; i < n; i++) {
      sims[i]->stream = own[i];
      sims[i]->exportSolverField();  // on the member's own stream; the batch's has finished
    }
    size_t off = 0;
    auto f0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n; i++) {
      fp_sim *s = sims[i];
      s->readActivity(pinnedReadback + off);
      off += 4 + (size_t)s->capacity;
      s->time += dt;
      s->steps++;
      s->stepped = true;
    }
    phaseMs[4] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - f0).count();
    lastMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  // The solve's structure, the same for every member.
  struct SolveShape {
    size_t args;
    int brickLevels, bottom;
    bool smallBottom, bottom64;
    bool anyFused, anyLarge;  // members whose tile levels run fused / level by level
  };

  // One V-cycle over the batch (PressureSolver.ts:303-363) in storage St; returns the level
  // flips (bit L: level L's iterate is in q). With `dry`, only the flips.
  template <class St> int vcycle(const SolveShape &c, bool dry) {
    using Args = MGArgsT<typename St::T>;
    int f = 0;
    auto go = [&](auto kernel, unsigned threads, size_t starts, unsigned g, int code) {
      if (!dry) launch<Args>(kernel, threads, c.args, starts, g, code);
    };
    auto relax = [&](int L) {
      if (L < c.brickLevels) go(k_mg_relax_bricks_batch<St>, 256, levelPrefix(L), wide(), L | f << 8);
      else go(k_mg_relax_tiles_batch<St>, 64, levelPrefix(L), wide(), L | f << 8);
      f ^= 1 << L;
    };
    std::function<void(int)> cycle = [&](int L) {
      if (L == c.bottom) {
        if (c.bottom64) go(k_mg_bottom64_batch<St>, 1024, levelPrefix(MAX_LEVELS), perSim(), f << 8);
        else if (c.smallBottom) go(k_mg_bottom_batch<St>, 1024, levelPrefix(MAX_LEVELS), perSim(), f << 8);
        else
          for (int i = 0; i < BOTTOM_ITERATIONS; i++) relax(L);
        return;
      }
      if (L == c.brickLevels) {
        // The first tile level (fp_sim::cycle): members of at most TC_MAX tile slots run the
        // levels below it in one workgroup each, the others level by level (the per-level
        // kernels have no blocks for the former).
        go(k_mg_tiles_down_batch<St>, 256, levelPrefix(L), wide(), f);
        f ^= 1 << L;
        if (c.anyFused)
          go(k_mg_coarse_cycle_batch<St>, TC_LANES, levelPrefix(MAX_LEVELS), perSim(), f | (c.bottom64 ? 1 << 16 : 0));
        if (c.anyLarge) cycle(L + 1);
        go(k_mg_tiles_up_batch<St>, 256, levelPrefix(L), wide(), f);
        f ^= 1 << L;
        return;
      }
      if (L >= 1 && L == c.brickLevels - 1) {
        // The last brick level (2^3 cells a brick) by tile of bricks and halo (fp_sim::cycle),
        // one workgroup per tile slot of each member.
        go(k_mg_bricks_down_batch<St>, 256, levelPrefix(c.brickLevels), wide(), f);
        f ^= 1 << L;
        cycle(L + 1);
        go(k_mg_bricks_up_batch<St>, 256, levelPrefix(c.brickLevels), wide(), f);
        f ^= 1 << L;
        return;
      }
      if (L == 0) go(k_mg_init_batch<St>, 256, levelPrefix(0), wide(), f << 8);
      else {
        relax(L);
        relax(L);
      }
      if (L + 1 < c.brickLevels) go(k_mg_restrict_bricks_batch<St>, 256, levelPrefix(L + 1), wide(), (L + 1) | f << 8);
      else go(k_mg_restrict_tiles_batch<St>, 64, levelPrefix(L + 1), wide(), (L + 1) | f << 8);
      cycle(L + 1);
      if (L < c.brickLevels) go(k_mg_prolong_bricks_batch<St>, 256, levelPrefix(L), wide(), L | f << 8);
      else go(k_mg_prolong_tiles_batch<St>, 64, levelPrefix(L), wide(), L | f << 8);
      f ^= 1 << L;
      relax(L);
      relax(L);
      relax(L);
    };
    cycle(0);
    return f;
  }

  // Defect correction over the batch: every simulation runs the loop of
  // fp_sim::correct, for up to `maxCycles` cycles; one that has finished (its own
  // tolerance met or its own cycle budget spent) is skipped by the rest, on the device,
  // so no launch waits for the host.
  template <class St> void levelPrefixes(const SolveShape &c) {
    k_level_prefix_batch<St><<<1, 1024, 0, stream>>>(
        reinterpret_cast<const MGArgsT<typename St::T> *>(devArgs.p + c.args), lay.cap,
        reinterpret_cast<int *>(devArgs.p + lay.levels));
  }
  // Defect correction over the batch, in the order of fp_sim::correct: after `it` V-cycles
  // the residual and its norms decide, on the device, whether a member has finished (its
  // own tolerance met or its own cycle budget spent); finished members drop out of every
  // later launch through the level prefixes. correctStart runs the loop's start and its
  // first cycles; correctCycles the cycles after `first - 1` up to `last`.
  template <class St> void correctStart(const SolveShape &c, int last) {
    using Args = MGArgsT<typename St::T>;
    launch<Args>(k_dc_start_batch<St>, 256, c.args, lay.slots, wide(), 0);
    levelPrefixes<St>(c);
    launch<Args>(k_dc_residual_batch<St>, 256, c.args, levelPrefix(0), wide(), 0);
    launch<Args>(k_dc_norms_batch<St>, 256, c.args, levelPrefix(MAX_LEVELS), perSim(), 0);
    levelPrefixes<St>(c);
    correctCycles<St>(c, 1, last);
  }
  template <class St> void correctCycles(const SolveShape &c, int first, int last) {
    using Args = MGArgsT<typename St::T>;
    for (int it = first; it <= last; it++) {
      launch<Args>(k_dc_scale_batch<St>, 256, c.args, levelPrefix(0), wide(), 0);
      const int flips = vcycle<St>(c, false);
      launch<Args>(k_dc_update_batch<St>, 256, c.args, levelPrefix(0), wide(), flips << 8);
      launch<Args>(k_dc_residual_batch<St>, 256, c.args, levelPrefix(0), wide(), 0);
      launch<Args>(k_dc_norms_batch<St>, 256, c.args, levelPrefix(MAX_LEVELS), perSim(), it);
      levelPrefixes<St>(c);  // members that finished drop out
    }
  }
  template <class F> void forMode(int mode, F f) {
    if (mode == 1) f(StoreF32{});
    else if (mode == 2) f(StoreBF16{});
    else f(StoreF16{});
  }

  // A step's kernels up to the pressure solve and, with the V-cycle policy, to its end; with
  // defect correction, the loop's start and its cycles up to `last`.
  void launchHead(const SolveShape &shape, int mode, int last, int zchunks, int fchunks) {
    // Allocation (fluid-allocation.wgsl).
    launch<AllocArgs>(k_build_bricks_batch, 256, lay.alloc, lay.blocks, wide());
    launch<AllocArgs>(k_scan_blocks_batch, 1024, lay.alloc, lay.one, perSim());
    launch<AllocArgs>(k_free_bricks_batch, 256, lay.alloc, lay.blocks, wide());
    launch<AllocArgs>(k_allocate_bricks_batch, 256, lay.alloc, lay.blocks, wide());
    launch<AllocArgs>(k_link_bricks_batch, 256, lay.alloc, lay.link, wide());
    launch<AllocArgs>(k_build_solids_batch, 64, lay.alloc, lay.slots, wide());
    launch<AllocArgs>(k_zero_fresh_batch, 256, lay.alloc, lay.zero, wide(), zchunks);
    // Transport, combustion, forces.
    launch<StepArgs>(k_advect_velocity_batch, vLanes, lay.pre, lay.slots, wide());
    launch<StepArgs>(k_correct_velocity_batch, vLanes, lay.pre, lay.slots, wide());
    launch<StepArgs>(k_advect_scalars_batch, 256, lay.pre, lay.field, wide(), fchunks);
    launch<StepArgs>(k_evolve_fields_batch, 256, lay.pre, lay.field, wide(), fchunks);
    launch<StepArgs>(k_curl_batch, vLanes, lay.post, lay.slots, wide());
    launch<StepArgs>(k_cell_forces_batch, vLanes, lay.post, lay.slots, wide());
    launch<StepArgs>(k_forces_batch, vLanes, lay.post, lay.slots, wide());
    launch<StepArgs>(k_divergence_batch, vLanes, lay.post, lay.slots, wide());
    // Pressure: topology, flags, then one V-cycle or defect correction.
    launch<MGArgs>(k_topology_bricks_batch, vLanes, lay.mg, lay.slots, wide(), 0);
    if (shape.bottom >= shape.brickLevels) launch<MGArgs>(k_topology_tiles_batch, 128, lay.mg, lay.tiles, wide(), -1);
    launch<MGArgs>(k_topology_bricks_batch, vLanes, lay.mg, lay.slots2, wide(), 1);
    launch<MGArgs>(k_eflags_bricks_batch, vLanes, lay.mg, lay.slots, wide());
    if (shape.bottom >= shape.brickLevels) launch<MGArgs>(k_eflags_tiles_batch, 64, lay.mg, lay.tiles, wide());
    if (mode == 0) {
      levelPrefixes<StoreF32>(shape);
      vcycle<StoreF32>(shape, false);
      launchTail();
    } else {
      forMode(mode, [&](auto st) { correctStart<decltype(st)>(shape, last); });
    }
  }
  // The projection and the activity readback.
  void launchTail() {
    launch<StepArgs>(k_project_batch, vLanes, lay.post, lay.slots, wide());
    k_readback_batch<<<perSim(), 256, 0, stream>>>(reinterpret_cast<const ReadbackArgs *>(devArgs.p + lay.rb),
                                                   lay.cap);
  }

  // Runs `body`'s launches as a graph captured once per key, or directly.
  template <class F> void replay(const GraphKey &key, F body) {
    if (!graphs) {
      body();
      return;
    }
    auto it = execs.find(key);
    if (it == execs.end()) {
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      body();
      HIPCHECK(hipStreamEndCapture(stream, &g));
      hipGraphExec_t exec = nullptr;
      hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHECK(err);
      it = execs.emplace(key, exec).first;
    }
    HIPCHECK(hipGraphLaunch(it->second, stream));
  }

  // Defect correction runs HEAD_CYCLES cycles with the step, then CHUNK_CYCLES at a time while
  // any member is still running (one 4-byte readback between chunks), so a step costs the
  // cycles its slowest member needs rather than the largest budget.
  static constexpr int HEAD_CYCLES = 8, CHUNK_CYCLES = 2;

  void enqueue(fp_sim *const *sims, int n) {
    const auto e0 = std::chrono::steady_clock::now();
    fp_sim &s0 = *sims[0];
    const int nF = s0.nF, mode = s0.pressureMode;
    const int zchunks = std::max(1, nF / 1024), fchunks = std::max(1, nF / 256);
    std::vector<AllocArgs> alloc(n);
    std::vector<StepArgs> pre(n), post(n);
    std::vector<MGArgs> mg(n), solve32;
    std::vector<MGArgsT<uint16_t>> solveBF16;
    std::vector<MGArgsT<_Float16>> solveF16;
    std::vector<ReadbackArgs> rb(n);
    std::vector<int> cBlocks(n), cOne(n, 1), cLink(n), cSlots(n), cSlots2(n), cZero(n), cField(n), cTiles(n);
    size_t words = 0;
    bool smallBottom = true, anyFused = false, anyLarge = false;
    int maxCycles = 0;
    for (int i = 0; i < n; i++) {
      fp_sim &s = *sims[i];
      const int slotsGrid = (std::max(1, s.poolSlots) + 7) / 8 * 8;
      cBlocks[i] = (s.brickCount() + 255) / 256;
      cLink[i] = (slotsGrid + LINK_BRICKS - 1) / LINK_BRICKS;
      cSlots[i] = slotsGrid;
      cSlots2[i] = slotsGrid * 2;
      cZero[i] = slotsGrid * zchunks;
      cField[i] = slotsGrid * fchunks;
      cTiles[i] = s.capacity;
      smallBottom = smallBottom && s.capacity <= BOTTOM64_TILES;
      (s.tilesFused() ? anyFused : anyLarge) = true;
      alloc[i] = s.allocArgs();
      s.solidsStale = false;
      pre[i] = s.stepArgs();
      std::swap(s.F, s.Fn);
      post[i] = s.stepArgs();
      mg[i] = s.mgArgs();
      mg[i].topologyAll = s.geometryDirty || s.tilesMoved ? 1 : 0;
      mg[i].topologyTiles = mg[i].topologyAll;
      s.geometryDirty = false;
      s.tilesMoved = false;
      if (mode != 0) {
        s.ensureCorrectionBuffers();
        if (mode == 1) solve32.push_back(s.correctionArgs32(mg[i]));
        else if (mode == 2) solveBF16.push_back(s.correctionArgs16<uint16_t>(mg[i]));
        else solveF16.push_back(s.correctionArgs16<_Float16>(mg[i]));
        maxCycles = std::max(maxCycles, s.pressureMaxCycles);
      }
      words += 4 + (size_t)s.capacity;
    }
    reserve(n);
    SolveShape shape{};
    shape.args = mode == 0 ? lay.mg : lay.solve;
    shape.brickLevels = s0.brickLevels;
    shape.bottom = s0.bottom;
    shape.smallBottom = smallBottom;
    shape.bottom64 = mode != 0;
    shape.anyFused = anyFused;
    shape.anyLarge = anyLarge;
    const int tm = (anyFused ? 1 : 0) | (anyLarge ? 2 : 0);
    // Where level 0's final iterate lands (P0 or P1, the same for every simulation), from a
    // dry run of the V-cycle; defect correction leaves its solution in the pair x.
    const int flips = mode == 0 ? vcycle<StoreF32>(shape, true) : 0;
    for (int i = 0; i < n; i++) {
      post[i].pressure = mode != 0 ? sims[i]->dcX.p : (flips & 1) ? mg[i].lv[0].q : mg[i].lv[0].p;
      post[i].pressureLo = mode != 0 ? sims[i]->dcXlo.p : nullptr;
    }
    if (readback.n < words) readback.alloc(words + 1024);
    if (pinnedWords < words) {
      if (pinnedReadback) (void)hipHostFree(pinnedReadback);
      pinnedReadback = nullptr;
      pinnedWords = words + 1024;
      HIPCHECK(hipHostMalloc((void **)&pinnedReadback, pinnedWords * sizeof(int), hipHostMallocDefault));
    }
    size_t o = 0;
    for (int i = 0; i < n; i++) {
      fp_sim &s = *sims[i];
      rb[i] = ReadbackArgs{s.counts.p, s.poolState.p, s.activity.p + s.brickCount(), s.capacity, readback.p + o,
                           mode != 0 ? s.dcNorms.p : nullptr, reports.p + 2 * i};
      o += 4 + (size_t)s.capacity;
    }
    HIPCHECK(hipStreamSynchronize(stream));  // the previous step's upload has been read
    put(lay.alloc, alloc);
    put(lay.pre, pre);
    put(lay.post, post);
    put(lay.rb, rb);
    put(lay.mg, mg);
    if (mode == 1) put(lay.solve, solve32);
    else if (mode == 2) put(lay.solve, solveBF16);
    else if (mode == 3) put(lay.solve, solveF16);
    putPrefix(lay.blocks, cBlocks, 1);
    putPrefix(lay.one, cOne, 1);
    putPrefix(lay.link, cLink, 1);
    putPrefix(lay.slots, cSlots, 8);
    putPrefix(lay.slots2, cSlots2, 8);
    putPrefix(lay.zero, cZero, 8);
    putPrefix(lay.field, cField, 8);
    putPrefix(lay.tiles, cTiles, 8);
    auto l0 = std::chrono::steady_clock::now();
    phaseMs[1] = std::chrono::duration<double, std::milli>(l0 - e0).count();
    HIPCHECK(hipMemcpyAsync(devArgs.p, hostArgs, lay.bytes, hipMemcpyHostToDevice, stream));
    const int sb = smallBottom ? 1 : 0;
    if (mode == 0) {
      replay(GraphKey{0, 0, 0, shape.brickLevels, shape.bottom, sb, zchunks, fchunks, lay.cap, tm},
             [&] { launchHead(shape, 0, 0, zchunks, fchunks); });
    } else {
      int done = std::min(HEAD_CYCLES, maxCycles);
      replay(GraphKey{0, mode, done, shape.brickLevels, shape.bottom, sb, zchunks, fchunks, lay.cap, tm},
             [&] { launchHead(shape, mode, done, zchunks, fchunks); });
      const int *live = reinterpret_cast<const int *>(devArgs.p + levelPrefix(MAX_LEVELS)) + lay.cap;
      while (done < maxCycles) {
        HIPCHECK(hipMemcpyAsync(pinnedLive, live, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIPCHECK(hipStreamSynchronize(stream));
        if (*pinnedLive == 0) break;
        const int first = done + 1, last = std::min(done + CHUNK_CYCLES, maxCycles);
        replay(GraphKey{1, mode, first, last, shape.brickLevels, shape.bottom, sb, 0, lay.cap, tm}, [&] {
          forMode(mode, [&](auto st) { correctCycles<decltype(st)>(shape, first, last); });
        });
        done = last;
      }
      replay(GraphKey{2, 0, 0, 0, 0, 0, 0, 0, lay.cap, 0}, [&] { launchTail(); });
    }
    HIPCHECK(hipMemcpyAsync(pinnedReadback, readback.p, words * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (mode != 0)
      HIPCHECK(hipMemcpyAsync(pinnedReports, reports.p, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream));
    auto w0 = std::chrono::steady_clock::now();
    phaseMs[2] = std::chrono::duration<double, std::milli>(w0 - l0).count();
    HIPCHECK(hipStreamSynchronize(stream));
    phaseMs[3] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    HIPCHECK(hipGetLastError());
    for (int i = 0; i < n; i++) {
      fp_sim &s = *sims[i];
      s.latestPressure = const_cast<float *>(post[i].pressure);
      s.latestPressureLo = const_cast<float *>(post[i].pressureLo);
      s.lastCycles = mode != 0 ? (int)pinnedReports[2 * i + 1] : 1;
      s.lastResidual = mode != 0 ? pinnedReports[2 * i] : -1.0;
    }
  }
};

extern "C" {

const char *fp_last_error(void) { return g_error.c_str(); }
const char *fp_version(void) { return "firepro-mi355x 0.1 (gfx950)"; }

int fp_sim_create(const fp_grid *grid, fp_sim **out) {
  if (!grid || !out) {
    g_error = "fp_sim_create: null argument";
    return FP_EINVAL;
  }
  *out = nullptr;
  fp_sim *sim = new fp_sim();
  int rc = guard([&] { sim->create(*grid); });
  if (rc != FP_OK) {
    sim->destroy();
    delete sim;
    return rc;
  }
  *out = sim;
  return FP_OK;
}

int fp_sim_step(fp_sim *sim, const fp_frame *frame, double dt) {
  if (!sim || !frame) {
    g_error = "fp_sim_step: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->step(*frame, dt); });
}

int fp_sim_reset(fp_sim *sim) {
  if (!sim) {
    g_error = "fp_sim_reset: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->reset(); });
}

void fp_sim_destroy(fp_sim *sim) {
  if (!sim) return;
  sim->destroy();
  delete sim;
}

int fp_sim_set_cutoff(fp_sim *sim, double cutoff) {
  if (!sim || !std::isfinite(cutoff)) {
    g_error = "cutoff must be a finite number.";
    return FP_EINVAL;
  }
  sim->cutoff = cutoff;
  return FP_OK;
}
int fp_sim_set_budget(fp_sim *sim, double maxVoxels) {
  if (!sim || !(maxVoxels > 0)) {
    g_error = "voxel budget must be positive.";
    return FP_EINVAL;
  }
  sim->voxelBudget = maxVoxels;
  return FP_OK;
}
int fp_sim_set_ground(fp_sim *sim, int32_t ground) {
  if (!sim) {
    g_error = "fp_sim_set_ground: null argument";
    return FP_EINVAL;
  }
  sim->ground = ground != 0;
  return FP_OK;
}

int fp_sim_stats(const fp_sim *sim, fp_stats *out) {
  if (!sim || !out) {
    g_error = "fp_sim_stats: null argument";
    return FP_EINVAL;
  }
  out->time = sim->time;
  out->steps = sim->steps;
  out->active_bricks = sim->activeBricks;
  out->active_voxels = sim->activeBricks * (int64_t)sim->nF;
  out->pool_slots = sim->poolSlots;
  out->used_slots = sim->usedSlots;
  out->pool_limited = sim->poolLimited;
  out->tile_capacity = sim->capacity;
  out->tile_count = sim->dir.count();
  out->pool_side = sim->side;
  out->memory_bytes = (int64_t)sim->memoryBytes();
  out->mesh_memory_bytes = (int64_t)sim->meshMemoryBytes();
  out->force_memory_bytes = (int64_t)sim->forceMemoryBytes();
  return FP_OK;
}

int fp_sim_show_solver_field(fp_sim *sim, int32_t field) {
  if (!sim) {
    g_error = "fp_sim_show_solver_field: null argument";
    return FP_EINVAL;
  }
  if (field != FP_SOLVER_FIELD_NONE && field != FP_SOLVER_FIELD_PRESSURE && field != FP_SOLVER_FIELD_DIVERGENCE) {
    g_error = "solver field must be pressure, divergence or none.";
    return FP_EINVAL;
  }
  return guard([&] {
    sim->drain();
    sim->solverField = field;
    if (!field) {
      sim->solverView.release();
      return;
    }
    if (sim->stepped) sim->exportSolverField();  // now, as well as after every step
  });
}

int fp_sim_read_solver_field(fp_sim *sim, int32_t *ids, float *values, int32_t capacity) {
  if (!sim) {
    g_error = "fp_sim_read_solver_field: null argument";
    return FP_EINVAL;
  }
  int result = 0;
  int rc = guard([&] {
    sim->drain();
    if (!sim->solverField) throw std::logic_error("no solver field is shown (fp_sim_show_solver_field).");
    int count = 0;
    if (sim->stepped && sim->solverView.p) HIPCHECK(hipMemcpy(&count, sim->counts.p, sizeof(int), hipMemcpyDeviceToHost));
    if (count > capacity) throw InvalidArgument("solver field buffers hold fewer bricks than are active.");
    std::vector<int32_t> act(count), entries(sim->brickCount());
    if (count) HIPCHECK(hipMemcpy(act.data(), sim->active.p, count * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(entries.data(), sim->entries.p, entries.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    const int nV = sim->nV;
    for (int i = 0; i < count; i++) {
      if (ids) ids[i] = act[i];
      if (values)
        HIPCHECK(hipMemcpy(values + (size_t)i * nV,
                           sim->solverView.p + (size_t)sim->pageToSlotHost(entries[act[i]] & 0xffffff) * nV,
                           nV * sizeof(float), hipMemcpyDeviceToHost));
    }
    result = count;
  });
  return rc != FP_OK ? rc : result;
}

int fp_sim_render_bricks(fp_sim *sim, fp_render_bricks *out) {
  if (!sim || !out) {
    g_error = "fp_sim_render_bricks: null argument";
    return FP_EINVAL;
  }
  return guard([&] {
    sim->drain();
    const auto &rb = sim->rboxes;
    out->version = rb.version;
    out->box_count = (int32_t)rb.boxes.size();
    out->tile_count = (int32_t)(rb.tiles.size() / 4);
    if (out->boxes)
      for (int i = 0; i < std::min(out->box_count, out->box_capacity); i++)
        for (int a = 0; a < 3; a++) {
          out->boxes[6 * i + a] = rb.boxes[i].lo[a];
          out->boxes[6 * i + 3 + a] = rb.boxes[i].hi[a];
        }
    if (out->tiles)
      std::memcpy(out->tiles, rb.tiles.data(), (size_t)std::min(out->tile_count, out->tile_capacity) * 4 * sizeof(int32_t));
    int counts[8] = {0};
    if (sim->stepped) HIPCHECK(hipMemcpy(counts, sim->counts.p, sizeof(counts), hipMemcpyDeviceToHost));
    out->pages_device = sim->pages.p;
    out->active_device = sim->active.p;
    out->zeroed_device = sim->zeroList.p;
    out->brick_count = sim->brickCount();
    out->active_count = counts[0];
    out->zeroed_count = counts[1] + counts[5];
  });
}

int fp_sim_read_bricks(fp_sim *sim, double *mins, int32_t *count) {
  if (!sim || !count) {
    g_error = "fp_sim_read_bricks: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->drain();
    int n = sim->brickCount();
    std::vector<int32_t> pages(n);
    HIPCHECK(hipMemcpy(pages.data(), sim->pages.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    double size = sim->B * sim->voxel;
    int written = 0;
    for (int id = 0; id < n; id++) {
      int t = id / 64;
      if (pages[id] < 0 || t >= (int)sim->dir.coords.size() || !sim->dir.coords[t][3]) continue;
      int l = id % 64;
      int loc[3] = {l & 3, (l >> 2) & 3, l >> 4};
      if (mins && written < *count)
        for (int a = 0; a < 3; a++) mins[written * 3 + a] = (sim->dir.coords[t][a] * 4 + loc[a]) * size;
      written++;
    }
    *count = written;
  });
}

int fp_sim_read_tables(fp_sim *sim, fp_tables *t) {
  if (!sim || !t) {
    g_error = "fp_sim_read_tables: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->drain();
    int n = sim->brickCount();
    t->brick_count = n;
    t->pool_slots = sim->poolSlots;
    t->free_count = sim->freeCount;
    t->failures = sim->lastFailures;
    int counts[8];
    HIPCHECK(hipMemcpy(counts, sim->counts.p, sizeof(counts), hipMemcpyDeviceToHost));
    int activeCount = sim->stepped ? counts[0] : 0;
    int zeroCount = sim->stepped ? counts[1] + counts[5] : 0;
    std::vector<int32_t> entries(n);
    HIPCHECK(hipMemcpy(entries.data(), sim->entries.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (t->pages) HIPCHECK(hipMemcpy(t->pages, sim->pages.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (t->entries) std::memcpy(t->entries, entries.data(), n * sizeof(int32_t));
    if (t->records) {
      HIPCHECK(hipMemcpy(t->records, sim->records.p, (size_t)n * 28 * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; i++)
        if (entries[i] < 0) std::memset(t->records + (size_t)i * 28, 0, 28 * sizeof(int32_t));
    }
    if (t->free_slots)
      HIPCHECK(hipMemcpy(t->free_slots, sim->freeSlots.p, sim->poolSlots * sizeof(uint32_t), hipMemcpyDeviceToHost));
    t->active_count = activeCount;
    if (t->active && activeCount)
      HIPCHECK(hipMemcpy(t->active, sim->active.p, activeCount * sizeof(int32_t), hipMemcpyDeviceToHost));
    t->zeroed_count = zeroCount;
    if (t->zeroed && zeroCount) {
      std::vector<std::pair<int32_t, uint32_t>> z(zeroCount);
      std::vector<int32_t> raw(2 * zeroCount);
      HIPCHECK(hipMemcpy(raw.data(), sim->zeroList.p, raw.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (int i = 0; i < zeroCount; i++) z[i] = {raw[2 * i], (uint32_t)raw[2 * i + 1]};
      std::sort(z.begin(), z.end());
      for (int i = 0; i < zeroCount; i++) {
        t->zeroed[2 * i] = z[i].first;
        t->zeroed[2 * i + 1] = (int32_t)z[i].second;
      }
    }
  });
}

int fp_sim_read_state(fp_sim *sim, int32_t *ids, float *fields, float *velocity, float *pressure, float *expansion,
                      int32_t capacity) {
  if (!sim) {
    g_error = "fp_sim_read_state: null argument";
    return FP_EINVAL;
  }
  int result = 0;
  int rc = guard([&] { sim->drain();
    int count = 0;
    if (sim->stepped) HIPCHECK(hipMemcpy(&count, sim->counts.p, sizeof(int), hipMemcpyDeviceToHost));
    if (count > capacity) throw InvalidArgument("state buffers hold fewer bricks than are active.");
    std::vector<int32_t> act(count), entries(sim->brickCount());
    if (count) HIPCHECK(hipMemcpy(act.data(), sim->active.p, count * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(entries.data(), sim->entries.p, entries.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    const int nF = sim->nF, nV = sim->nV;
    for (int i = 0; i < count; i++) {
      int id = act[i];
      if (ids) ids[i] = id;
      int slot = (int)((((uint32_t)(entries[id] & 0xffffff) >> 16) * sim->side + (((uint32_t)entries[id] >> 8) & 255u)) *
                           sim->side +
                       ((uint32_t)entries[id] & 255u));
      if (fields)
        HIPCHECK(hipMemcpy(fields + (size_t)i * nF * 4, sim->F.p + (size_t)slot * nF, nF * sizeof(float4),
                           hipMemcpyDeviceToHost));
      if (expansion)
        HIPCHECK(hipMemcpy(expansion + (size_t)i * nF, sim->E.p + (size_t)slot * nF, nF * sizeof(float),
                           hipMemcpyDeviceToHost));
      if (velocity)
        HIPCHECK(hipMemcpy(velocity + (size_t)i * nV * 3, sim->V.p + (size_t)slot * nV * 3, nV * 3 * sizeof(float),
                           hipMemcpyDeviceToHost));
      if (pressure)
        HIPCHECK(hipMemcpy(pressure + (size_t)i * nV, sim->latestPressure + (size_t)slot * nV, nV * sizeof(float),
                           hipMemcpyDeviceToHost));
    }
    result = count;
  });
  return rc != FP_OK ? rc : result;
}

int fp_sim_write_state(fp_sim *sim, const float *fields, const float *velocity) {
  if (!sim) {
    g_error = "fp_sim_write_state: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->drain();
    int count = 0;
    if (sim->stepped) HIPCHECK(hipMemcpy(&count, sim->counts.p, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int32_t> act(count), entries(sim->brickCount());
    if (count) HIPCHECK(hipMemcpy(act.data(), sim->active.p, count * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(entries.data(), sim->entries.p, entries.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    const int nF = sim->nF, nV = sim->nV;
    for (int i = 0; i < count; i++) {
      uint32_t e = (uint32_t)entries[act[i]] & 0xffffffu;
      int slot = (int)(((e >> 16) * sim->side + ((e >> 8) & 255u)) * sim->side + (e & 255u));
      if (fields)
        HIPCHECK(hipMemcpy(sim->F.p + (size_t)slot * nF, fields + (size_t)i * nF * 4, nF * sizeof(float4),
                           hipMemcpyHostToDevice));
      if (velocity)
        HIPCHECK(hipMemcpy(sim->V.p + (size_t)slot * nV * 3, velocity + (size_t)i * nV * 3, nV * 3 * sizeof(float),
                           hipMemcpyHostToDevice));
    }
  });
}

int fp_sim_read_smoke(fp_sim *sim, int32_t *ids, float *density, int32_t capacity) {
  if (!sim) {
    g_error = "fp_sim_read_smoke: null argument";
    return FP_EINVAL;
  }
  int result = 0;
  int rc = guard([&] { sim->drain();
    if (sim->smokeDiv == 1) throw InvalidArgument("this simulation has no coarse smoke grid (smokeDivisor 1).");
    int count = 0;
    if (sim->stepped) HIPCHECK(hipMemcpy(&count, sim->counts.p, sizeof(int), hipMemcpyDeviceToHost));
    if (count > capacity) throw InvalidArgument("smoke buffers hold fewer bricks than are active.");
    std::vector<int32_t> act(count), entries(sim->brickCount());
    if (count) HIPCHECK(hipMemcpy(act.data(), sim->active.p, count * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(entries.data(), sim->entries.p, entries.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    const int nS = sim->nS;
    for (int i = 0; i < count; i++) {
      if (ids) ids[i] = act[i];
      if (density)
        HIPCHECK(hipMemcpy(density + (size_t)i * nS, sim->S.p + (size_t)sim->pageToSlotHost(entries[act[i]] & 0xffffff) * nS,
                           nS * sizeof(float), hipMemcpyDeviceToHost));
    }
    result = count;
  });
  return rc != FP_OK ? rc : result;
}

void *fp_sim_stream(fp_sim *sim) { return sim ? (void *)sim->stream : nullptr; }

int fp_sim_profile(fp_sim *sim, int32_t enable) {
  if (!sim) {
    g_error = "fp_sim_profile: null argument";
    return FP_EINVAL;
  }
  sim->profiling = enable != 0;
  sim->profileLevel = enable;
  sim->profile.clear();
  sim->profileNames.clear();
  return FP_OK;
}

int fp_sim_profile_read(fp_sim *sim, const char **names, double *ms, int32_t *count) {
  if (!sim || !count) {
    g_error = "fp_sim_profile_read: null argument";
    return FP_EINVAL;
  }
  int n = std::min<int>(*count, (int)sim->profileNames.size());
  for (int i = 0; i < n; i++) {
    const auto &e = sim->profile[sim->profileNames[i]];
    if (names) names[i] = sim->profileNames[i].c_str();
    if (ms) ms[i] = e.second ? e.first / e.second : 0.0;
  }
  *count = n;
  return FP_OK;
}

int fp_sim_scene_lights(fp_sim *sim, const fp_render_settings *settings, const double *anchors, int32_t count,
                        float *sums) {
  if (!sim || !settings || (count > 0 && (!anchors || !sums))) {
    g_error = "fp_sim_scene_lights: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->drain(); sim->sceneLights(*settings, anchors, count, sums); });
}

int fp_sim_scene_lights_noalloc(fp_sim *sim, const fp_render_settings *settings, const double *anchors,
                                int32_t count, float *sums) {
  if (!sim || !settings || (count > 0 && (!anchors || !sums))) {
    g_error = "fp_sim_scene_lights_noalloc: null argument";
    return FP_EINVAL;
  }
  return guard([&] {
    sim->drain();
    unsigned char *base = noallocScratch();
    sim->sceneLightsNoAlloc(*settings, anchors, count, sums, base);
  });
}

int fp_batch_create(fp_batch **out) {
  if (!out) {
    g_error = "fp_batch_create: null argument";
    return FP_EINVAL;
  }
  return guard([&] {
    std::unique_ptr<fp_batch> b(new fp_batch());
    b->create();
    *out = b.release();
  });
}

void fp_batch_destroy(fp_batch *batch) {
  if (!batch) return;
  batch->destroy();
  delete batch;
}

int fp_batch_step(fp_batch *batch, fp_sim *const *sims, const fp_frame *frames, int32_t count, double dt) {
  if (!batch || (count > 0 && (!sims || !frames)) || count < 0) {
    g_error = "fp_batch_step: null argument";
    return FP_EINVAL;
  }
  return guard([&] {
    for (int i = 0; i < count; i++)
      if (sims[i]) sims[i]->drain();  // a single step's tail may still be queued
    batch->step(sims, frames, count, dt);
  });
}

int fp_batch_last_ms(const fp_batch *batch, double *ms) {
  if (!batch || !ms) {
    g_error = "fp_batch_last_ms: null argument";
    return FP_EINVAL;
  }
  ms[0] = batch->lastMs;
  return FP_OK;
}

int fp_batch_use_graphs(fp_batch *batch, int32_t enable) {
  if (!batch) {
    g_error = "fp_batch_use_graphs: null argument";
    return FP_EINVAL;
  }
  return guard([&] {
    HIPCHECK(hipStreamSynchronize(batch->stream));
    batch->graphs = enable != 0;
  });
}

int fp_batch_phases(const fp_batch *batch, double *ms) {
  if (!batch || !ms) {
    g_error = "fp_batch_phases: null argument";
    return FP_EINVAL;
  }
  for (int i = 0; i < 5; i++) ms[i] = batch->phaseMs[i];
  return FP_OK;
}

int fp_sim_set_pressure(fp_sim *sim, int32_t mode, double tolerance, int32_t max_cycles) {
  if (!sim) {
    g_error = "fp_sim_set_pressure: null argument";
    return FP_EINVAL;
  }
  if (mode < FP_PRESSURE_VCYCLE || mode > FP_PRESSURE_FP16) {
    g_error = "pressure mode must be one of FP_PRESSURE_VCYCLE, _FP32, _BF16, _FP16.";
    return FP_EINVAL;
  }
  if (!(tolerance >= 0) || max_cycles < 1 || max_cycles > 1000) {
    g_error = "pressure tolerance must be >= 0 and max_cycles from 1 to 1000.";
    return FP_EINVAL;
  }
  sim->pressureMode = mode;
  sim->pressureTol = tolerance;
  sim->pressureMaxCycles = max_cycles;
  return FP_OK;
}

int fp_sim_pressure_report(const fp_sim *sim, int32_t *cycles, double *relative_residual) {
  if (!sim) {
    g_error = "fp_sim_pressure_report: null argument";
    return FP_EINVAL;
  }
  if (cycles) *cycles = sim->lastCycles;
  if (relative_residual) *relative_residual = sim->lastResidual;
  return FP_OK;
}

int fp_sim_render(fp_sim *sim, const fp_camera *camera, const fp_render_settings *settings, float *rgba) {
  if (!sim || !camera || !settings || !rgba) {
    g_error = "fp_sim_render: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->drain(); sim->renderImage(*camera, *settings, rgba, false); });
}

int fp_sim_render_device(fp_sim *sim, const fp_camera *camera, const fp_render_settings *settings,
                         float *rgba_device) {
  if (!sim || !camera || !settings || !rgba_device) {
    g_error = "fp_sim_render_device: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->drain(); sim->renderImage(*camera, *settings, rgba_device, true); });
}

int fp_sim_render_stats(fp_sim *sim, uint64_t *counts) {
  if (!sim || !counts) {
    g_error = "fp_sim_render_stats: null argument";
    return FP_EINVAL;
  }
  return guard([&] {
    sim->drain();
    for (int k = 0; k < 4; k++) counts[k] = 0;
    if (sim->rc.sampleCounts.p)
      HIPCHECK(hipMemcpy(counts, sim->rc.sampleCounts.p, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost));
  });
}

int fp_sim_read_render(fp_sim *sim, fp_render_tables *out) {
  if (!sim || !out) {
    g_error = "fp_sim_read_render: null argument";
    return FP_EINVAL;
  }
  return guard([&] { sim->drain(); sim->readRender(out); });
}

}  // extern "C"
