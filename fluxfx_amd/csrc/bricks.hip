// Sparse 8^3 / 16^3 brick pool: the host topology (brick coordinate -> slot,
// FluxFX_Source/native/bricks.hpp:13-94), source-driven active regions with
// predicted motion, halo and delayed deactivation (native/regions.hpp:7-52),
// capacity growth by doubling (docs/DENSITY_CAPACITY.md), and the device
// mirror: the coordinate table, the active list and six neighbour slots per
// brick (built on the GPU from the device copy of the hash table), one fp32
// value per voxel of every slot, dense <-> brick gather/scatter and the
// per-brick support scan of a dense field.
//
// Slots come from a LIFO free stack (0 first) and the active list is compact
// (swap-remove on deactivation), so slot ids, active order, neighbour ids and
// errors follow the reference's topology for any command sequence; the hash
// table itself is this file's own (packed 63-bit keys, splitmix finaliser,
// linear probing, tombstones reclaimed by an in-place rebuild).
#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

namespace fx {

constexpr int32_t kCoordMin = -1048576, kCoordMax = 1048575;
constexpr uint32_t kMaxBricks = 262144;
constexpr uint64_t kEmptyKey = ~uint64_t(0);
constexpr int32_t kEmptySlot = -1, kTombstone = -2;

struct BC {
  int32_t x, y, z;
};

__host__ __device__ inline uint64_t pack_key(int32_t x, int32_t y, int32_t z) {
  return uint64_t(uint32_t(x - kCoordMin)) | (uint64_t(uint32_t(y - kCoordMin)) << 21) |
         (uint64_t(uint32_t(z - kCoordMin)) << 42);
}

__host__ __device__ inline uint64_t mix_key(uint64_t k) {
  k ^= k >> 30;
  k *= 0xbf58476d1ce4e5b9ull;
  k ^= k >> 27;
  k *= 0x94d049bb133111ebull;
  return k ^ (k >> 31);
}

static void validate(BC c) {
  for (int32_t v : {c.x, c.y, c.z})
    if (v < kCoordMin || v > kCoordMax)
      throw std::invalid_argument("Brick coordinates must be in [-1048576,1048575]");
}

// Fixed-capacity topology; capacity grows only through grow().
class BrickTopo {
 public:
  explicit BrickTopo(uint32_t capacity) {
    if (!capacity || capacity > kMaxBricks) throw std::invalid_argument("Brick capacity must be 1..262144");
    size_table(capacity);
    coords_.resize(capacity);
    pos_.resize(capacity);
    free_.reserve(capacity);
    active_.reserve(capacity);
    for (uint32_t i = capacity; i > 0; --i) free_.push_back(i - 1);
  }

  uint32_t capacity() const { return uint32_t(coords_.size()); }
  const std::vector<uint32_t>& active() const { return active_; }
  BC coord(uint32_t slot) const {
    if (slot >= capacity()) throw std::out_of_range("Brick slot out of range");
    return coords_[slot];
  }
  const std::vector<uint64_t>& keys() const { return keys_; }
  const std::vector<int32_t>& vals() const { return vals_; }
  uint64_t metadata_bytes() const {
    return keys_.capacity() * 8 + vals_.capacity() * 4 + coords_.capacity() * sizeof(BC) +
           (pos_.capacity() + free_.capacity() + active_.capacity()) * 4;
  }

  int32_t lookup(BC c) const {
    validate(c);
    const size_t i = find(pack_key(c.x, c.y, c.z));
    return i < keys_.size() && vals_[i] >= 0 ? vals_[i] : -1;
  }

  uint32_t activate(BC c) {
    validate(c);
    const uint64_t k = pack_key(c.x, c.y, c.z);
    const size_t at = find(k);
    if (at < keys_.size() && vals_[at] >= 0) return uint32_t(vals_[at]);
    if (free_.empty()) throw std::runtime_error("Brick pool capacity exhausted");
    // first tombstone or empty entry on the probe path
    const size_t mask = keys_.size() - 1;
    size_t p = mix_key(k) & mask;
    while (vals_[p] >= 0) p = (p + 1) & mask;
    if (vals_[p] == kTombstone) --tombstones_;
    const uint32_t slot = free_.back();
    free_.pop_back();
    keys_[p] = k;
    vals_[p] = int32_t(slot);
    coords_[slot] = c;
    pos_[slot] = uint32_t(active_.size());
    active_.push_back(slot);
    fresh_.push_back(slot);
    return slot;
  }

  // slots activated since the last call (a slot may repeat): their payload
  // still holds whatever the slot held before and must read as zero
  std::vector<uint32_t> take_fresh() {
    std::vector<uint32_t> out;
    out.swap(fresh_);
    return out;
  }

  bool deactivate(BC c) {
    validate(c);
    const size_t i = find(pack_key(c.x, c.y, c.z));
    if (i >= keys_.size() || vals_[i] < 0) return false;
    const uint32_t slot = uint32_t(vals_[i]), at = pos_[slot], last = active_.back();
    active_[at] = last;
    pos_[last] = at;
    active_.pop_back();
    vals_[i] = kTombstone;
    ++tombstones_;
    free_.push_back(slot);
    if (tombstones_ * 4 > keys_.size()) rehash(keys_.size());
    return true;
  }

  void neighbors(BC c, int32_t out[6]) const {
    validate(c);
    const BC adj[6] = {{c.x - 1, c.y, c.z}, {c.x + 1, c.y, c.z}, {c.x, c.y - 1, c.z},
                       {c.x, c.y + 1, c.z}, {c.x, c.y, c.z - 1}, {c.x, c.y, c.z + 1}};
    for (int d = 0; d < 6; ++d) {
      // the one-past neighbours of the coordinate limits are never stored
      const size_t i = find(pack_key_wrapped(adj[d]));
      out[d] = i < keys_.size() && vals_[i] >= 0 ? vals_[i] : -1;
    }
  }

  // every bound, the capacity and the free slots are checked before any change
  void activate_box(BC lo, BC ext) {
    validate(lo);
    if (ext.x < 1 || ext.y < 1 || ext.z < 1 || ext.x > 4096 || ext.y > 4096 || ext.z > 4096)
      throw std::invalid_argument("Box extent must be 1..4096 per axis");
    validate({lo.x + ext.x - 1, lo.y + ext.y - 1, lo.z + ext.z - 1});
    if (uint64_t(ext.x) * ext.y * ext.z > capacity()) throw std::runtime_error("Box exceeds pool capacity");
    uint32_t needed = 0;
    for (int z = 0; z < ext.z; ++z)
      for (int y = 0; y < ext.y; ++y)
        for (int x = 0; x < ext.x; ++x) needed += lookup({lo.x + x, lo.y + y, lo.z + z}) < 0;
    if (needed > free_.size()) throw std::runtime_error("Brick pool capacity exhausted");
    for (int z = 0; z < ext.z; ++z)
      for (int y = 0; y < ext.y; ++y)
        for (int x = 0; x < ext.x; ++x) activate({lo.x + x, lo.y + y, lo.z + z});
  }

  uint32_t free_count() const { return uint32_t(free_.size()); }

  // Capacity becomes `capacity` (larger): slots and active order are kept; the
  // new slots are handed out, lowest first, after the existing free ones.
  void grow(uint32_t capacity) {
    const uint32_t old = this->capacity();
    if (capacity <= old) return;
    std::vector<uint32_t> fresh;
    fresh.reserve(capacity);
    for (uint32_t i = capacity; i > old; --i) fresh.push_back(i - 1);
    fresh.insert(fresh.end(), free_.begin(), free_.end());
    free_.swap(fresh);
    coords_.resize(capacity);
    pos_.resize(capacity);
    active_.reserve(capacity);
    size_t n = 2;
    while (n < size_t(capacity) * 2) n *= 2;
    rehash(n);
  }

 private:
  std::vector<uint64_t> keys_;
  std::vector<int32_t> vals_;
  std::vector<BC> coords_;
  std::vector<uint32_t> pos_, free_, active_, fresh_;
  size_t tombstones_ = 0;

  static uint64_t pack_key_wrapped(BC c) {
    if (c.x < kCoordMin || c.x > kCoordMax || c.y < kCoordMin || c.y > kCoordMax || c.z < kCoordMin ||
        c.z > kCoordMax)
      return kEmptyKey - 1;  // never stored
    return pack_key(c.x, c.y, c.z);
  }

  void size_table(uint32_t capacity) {
    size_t n = 2;
    while (n < size_t(capacity) * 2) n *= 2;
    keys_.assign(n, kEmptyKey);
    vals_.assign(n, kEmptySlot);
    tombstones_ = 0;
  }

  size_t find(uint64_t k) const {
    const size_t mask = keys_.size() - 1;
    size_t p = mix_key(k) & mask;
    for (size_t n = 0; n < keys_.size(); ++n, p = (p + 1) & mask) {
      if (vals_[p] == kEmptySlot) return p;
      if (vals_[p] >= 0 && keys_[p] == k) return p;
    }
    return keys_.size();
  }

  void rehash(size_t n) {
    std::vector<uint64_t> keys(n, kEmptyKey);
    std::vector<int32_t> vals(n, kEmptySlot);
    const size_t mask = n - 1;
    for (size_t i = 0; i < keys_.size(); ++i) {
      if (vals_[i] < 0) continue;
      size_t p = mix_key(keys_[i]) & mask;
      while (vals[p] >= 0) p = (p + 1) & mask;
      keys[p] = keys_[i];
      vals[p] = vals_[i];
    }
    keys_.swap(keys);
    vals_.swap(vals);
    tombstones_ = 0;
  }
};

// Source-driven active regions (one brick-domain scratch mask, per-slot ages).
struct RegionSrc {
  double x, y, z, radius, vx, vy, vz;
};

class Regions {
 public:
  explicit Regions(uint32_t capacity) : wanted_(64 * 64 * 64), ages_(capacity, 0) {}
  uint32_t age(uint32_t slot) const {
    if (slot >= ages_.size()) throw std::out_of_range("Brick slot out of range");
    return ages_[slot];
  }
  void grow(uint32_t capacity) {
    if (capacity > ages_.size()) ages_.resize(capacity, 0);
  }
  uint64_t bytes() const { return wanted_.capacity() + ages_.capacity() * 4; }

  void update(BrickTopo& t, const std::vector<RegionSrc>& sources, int32_t grid, double dt, int32_t linger,
              int32_t halo) {
    if (grid < 1 || grid > 64 || !std::isfinite(dt) || dt <= 0 || dt > 1 || linger < 1 || linger > 1000 ||
        halo < 0 || halo > 4 || sources.size() > 64)
      throw std::invalid_argument(
          "Region settings: grid 1..64, dt (0,1], linger 1..1000, halo 0..4, sources <=64");
    std::fill(wanted_.begin(), wanted_.end(), 0);
    uint32_t count = 0;
    const double g = grid;
    for (const RegionSrc& s : sources) {
      for (double v : {s.x, s.y, s.z, s.radius, s.vx, s.vy, s.vz})
        if (!std::isfinite(v) || std::abs(v) > 1e6)
          throw std::invalid_argument("Region source must be finite and bounded");
      if (s.radius < 0) throw std::invalid_argument("Source radius must be nonnegative");
      const double c[3] = {s.x, s.y, s.z}, v[3] = {s.vx, s.vy, s.vz};
      int lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        // the sphere swept over dt, widened by the halo and clipped to the domain
        const double end = c[a] + v[a] * dt;
        lo[a] = int(std::floor(std::clamp((std::min(c[a], end) - s.radius) * g - double(halo), 0.0, g)));
        hi[a] = int(std::ceil(std::clamp((std::max(c[a], end) + s.radius) * g + double(halo), 0.0, g)));
      }
      for (int z = lo[2]; z < hi[2]; ++z)
        for (int y = lo[1]; y < hi[1]; ++y)
          for (int x = lo[0]; x < hi[0]; ++x) {
            uint8_t& w = wanted_[(size_t(z) * grid + y) * grid + x];
            if (!w) {
              w = 1;
              ++count;
            }
          }
    }
    auto desired = [&](BC c) {
      return c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x < grid && c.y < grid && c.z < grid &&
             wanted_[(size_t(c.z) * grid + c.y) * grid + c.x];
    };
    uint32_t retained = 0;
    for (uint32_t slot : t.active())
      if (!desired(t.coord(slot)) && ages_[slot] + 1 < uint32_t(linger)) ++retained;
    if (uint64_t(count) + retained > t.capacity())
      throw std::runtime_error(
          "Active region plus retained bricks exceeds pool capacity; increase pool or reduce halo");
    // nothing (ages included) changes before the capacity check has passed
    size_t i = 0;
    while (i < t.active().size()) {
      const uint32_t slot = t.active()[i];
      const BC c = t.coord(slot);
      if (desired(c)) {
        ages_[slot] = 0;
        ++i;
      } else if (++ages_[slot] >= uint32_t(linger)) {
        t.deactivate(c);  // the last active slot moves into position i
      } else {
        ++i;
      }
    }
    for (int z = 0; z < grid; ++z)
      for (int y = 0; y < grid; ++y)
        for (int x = 0; x < grid; ++x)
          if (wanted_[(size_t(z) * grid + y) * grid + x]) ages_[t.activate({x, y, z})] = 0;
  }

 private:
  std::vector<uint8_t> wanted_;
  std::vector<uint32_t> ages_;
};

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------
__device__ inline int32_t dev_lookup(const uint64_t* __restrict__ keys, const int32_t* __restrict__ vals,
                                     uint32_t mask, int32_t x, int32_t y, int32_t z) {
  if (x < kCoordMin || x > kCoordMax || y < kCoordMin || y > kCoordMax || z < kCoordMin || z > kCoordMax) return -1;
  const uint64_t k = pack_key(x, y, z);
  uint32_t p = uint32_t(mix_key(k)) & mask;
  for (uint32_t n = 0; n <= mask; ++n, p = (p + 1) & mask) {
    const int32_t v = vals[p];
    if (v == kEmptySlot) return -1;
    if (v >= 0 && keys[p] == k) return v;
  }
  return -1;
}

// one thread per (active brick, face): nbr[slot*6 + d] = slot of the face
// neighbour (-X, +X, -Y, +Y, -Z, +Z) or -1
__global__ void k_brick_neighbors(int n_active, const uint32_t* __restrict__ active, const int4* __restrict__ coords,
                                  const uint64_t* __restrict__ keys, const int32_t* __restrict__ vals, uint32_t mask,
                                  int32_t* __restrict__ nbr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_active * 6) return;
  const int b = t / 6, d = t - b * 6;
  const uint32_t slot = active[b];
  const int4 c = coords[slot];
  const int axis = d >> 1, step = (d & 1) ? 1 : -1;
  nbr[size_t(slot) * 6 + d] = dev_lookup(keys, vals, mask, c.x + (axis == 0 ? step : 0),
                                         c.y + (axis == 1 ? step : 0), c.z + (axis == 2 ? step : 0));
}

__global__ void k_fill_i32(int64_t n, int32_t* __restrict__ a, int32_t v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    a[i] = v;
}

// Dense <-> brick copies: one block per active brick, one voxel per thread
// and pass (side^3 voxels, x fastest, as the dense grid); voxels outside the
// dense grid read 0 and are not written.
template <bool GATHER>
__global__ __launch_bounds__(256) void k_brick_copy(int side, const uint32_t* __restrict__ active,
                                                    const int4* __restrict__ coords, int nx, int ny, int nz,
                                                    float* __restrict__ dense, float* __restrict__ field) {
  const uint32_t slot = active[blockIdx.x];
  const int4 c = coords[slot];
  const int vox = side * side * side;
  float* __restrict__ brick = field + size_t(slot) * vox;
  for (int v = threadIdx.x; v < vox; v += blockDim.x) {
    const int lx = v % side, ly = (v / side) % side, lz = v / (side * side);
    const int64_t x = int64_t(c.x) * side + lx, y = int64_t(c.y) * side + ly, z = int64_t(c.z) * side + lz;
    const bool in = x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
    const int64_t i = in ? (z * ny + y) * nx + x : 0;
    if (GATHER)
      brick[v] = in ? dense[i] : 0.0f;
    else if (in)
      dense[i] = brick[v];
  }
}

// flags[b] = 1 when any voxel of dense brick b (side^3, clipped to the grid)
// has |value| > threshold; bricks enumerated x fastest over ceil(n/side).
// One block per brick row segment: a wave covers 64 consecutive x voxels.
__global__ __launch_bounds__(256) void k_brick_support(int side, int nx, int ny, int nz, int bx, int by, int bz,
                                                       const float* __restrict__ dense, float threshold,
                                                       uint8_t* __restrict__ flags) {
  const int b = blockIdx.x;
  const int cx = b % bx, cy = (b / bx) % by, cz = b / (bx * by);
  const int vox = side * side * side;
  bool any = false;
  for (int v = threadIdx.x; v < vox; v += blockDim.x) {
    const int x = cx * side + v % side, y = cy * side + (v / side) % side, z = cz * side + v / (side * side);
    if (x < nx && y < ny && z < nz) any |= fabsf(dense[(int64_t(z) * ny + y) * nx + x]) > threshold;
  }
  const int found = __syncthreads_or(any ? 1 : 0);
  if (threadIdx.x == 0) flags[b] = found ? 1 : 0;
}

// Row form of the support scan (nx % 4 == 0, 16-byte aligned field): one
// thread per 4 consecutive x voxels of a (y, z) row, so the field streams
// through once in 16-byte loads; a hit stores 1 into its brick's flag (all
// stores of a brick write the same value), the flags having been cleared.
__global__ __launch_bounds__(256) void k_brick_support_rows(int side, int nx, int ny, int bx, int by, int64_t nq,
                                                            const float4* __restrict__ dense, float threshold,
                                                            uint8_t* __restrict__ flags) {
  const int qx = nx >> 2;
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nq; q += int64_t(gridDim.x) * blockDim.x) {
    const float4 v = dense[q];
    const bool hit = fabsf(v.x) > threshold || fabsf(v.y) > threshold || fabsf(v.z) > threshold ||
                     fabsf(v.w) > threshold;
    if (hit) {
      const int64_t row = q / qx;
      const int x = int(q - row * qx) << 2, y = int(row % ny), z = int(row / ny);
      flags[(int64_t(z / side) * by + y / side) * bx + x / side] = 1;
    }
  }
}

// one block per listed slot: zero that brick's payload
__global__ void k_brick_clear(int vox, const uint32_t* __restrict__ slots, float* __restrict__ field) {
  float* brick = field + size_t(slots[blockIdx.x]) * vox;
  for (int i = threadIdx.x; i < vox; i += blockDim.x) brick[i] = 0.0f;
}

__global__ void k_fill_u8(int64_t n, uint8_t* __restrict__ a) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    a[i] = 0;
}

// 16-byte form of the dense <-> brick copies (nx % 4 == 0, aligned): one
// float4 of a brick row per thread; a quad lies wholly inside or outside the grid.
template <bool GATHER>
__global__ __launch_bounds__(256) void k_brick_copy4(int side, const uint32_t* __restrict__ active,
                                                     const int4* __restrict__ coords, int nx, int ny, int nz,
                                                     float4* __restrict__ dense, float4* __restrict__ field) {
  const uint32_t slot = active[blockIdx.x];
  const int4 c = coords[slot];
  const int q4 = side >> 2, nq = q4 * side * side;
  float4* __restrict__ brick = field + size_t(slot) * nq;
  for (int v = threadIdx.x; v < nq; v += blockDim.x) {
    const int lx = (v % q4) << 2, ly = (v / q4) % side, lz = v / (q4 * side);
    const int64_t x = int64_t(c.x) * side + lx, y = int64_t(c.y) * side + ly, z = int64_t(c.z) * side + lz;
    const bool in = x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
    const int64_t i = in ? ((z * ny + y) * nx + x) >> 2 : 0;
    if (GATHER)
      brick[v] = in ? dense[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    else if (in)
      dense[i] = brick[v];
  }
}

// Transport-aware preflight (native/projection.mm:126-138 requiredDensity,
// with sampleFaceField/sampleVelocity of :101-103 and emission of :115-121):
// a brick of the cubic n^3 domain is flagged when any of its cells is
// emitted into, or the RK2 back-trace of the cell centre (velocity in cells
// per step, MAC faces) lands where one of the 8 donor cells holds density or
// is emitted into.  Evaluated in fp32 in the reference's operation order with
// contraction off, so the float32 oracle reproduces every flag.
struct ReqArgs {
  int n, side, grid;
  const float* density;  // n^3
  const float* u;        // (n+1) n n
  const float* v;        // n (n+1) n
  const float* w;        // n n (n+1)
  float cx, cy, cz, radius, rate;
};

__device__ inline float face_at(const ReqArgs& A, int a, int x, int y, int z) {
  const int sx = A.n + (a == 0), sy = A.n + (a == 1), sz = A.n + (a == 2);
  if (x < 0 || y < 0 || z < 0 || x >= sx || y >= sy || z >= sz) return 0.0f;
  const float* f = a == 0 ? A.u : (a == 1 ? A.v : A.w);
  return f[(int64_t(z) * sy + y) * sx + x];
}

__device__ inline float sample_face_cells(const ReqArgs& A, int a, float px, float py, float pz) {
#pragma clang fp contract(off)
  const float qx = fminf(fmaxf(px - (a == 0 ? 0.0f : 0.5f), 0.0f), float(A.n + (a == 0) - 1));
  const float qy = fminf(fmaxf(py - (a == 1 ? 0.0f : 0.5f), 0.0f), float(A.n + (a == 1) - 1));
  const float qz = fminf(fmaxf(pz - (a == 2 ? 0.0f : 0.5f), 0.0f), float(A.n + (a == 2) - 1));
  const float bx = floorf(qx), by = floorf(qy), bz = floorf(qz);
  const float fx = qx - bx, fy = qy - by, fz = qz - bz;
  float acc = 0.0f;
  for (int z = 0; z < 2; ++z)
    for (int y = 0; y < 2; ++y)
      for (int x = 0; x < 2; ++x) {
        const float wx = x ? fx : 1.0f - fx, wy = y ? fy : 1.0f - fy, wz = z ? fz : 1.0f - fz;
        const float t = wx * wy;
        const float wgt = t * wz;
        const float val = face_at(A, a, int(bx) + x, int(by) + y, int(bz) + z);
        const float term = wgt * val;
        acc = acc + term;
      }
  return acc;
}

__device__ inline bool emitted(const ReqArgs& A, int x, int y, int z) {
#pragma clang fp contract(off)
  const float dx = (float(x) + 0.5f - A.cx) / A.radius, dy = (float(y) + 0.5f - A.cy) / A.radius,
              dz = (float(z) + 0.5f - A.cz) / A.radius;
  const float dd = dx * dx + dy * dy + dz * dz;
  return A.rate * fmaxf(0.0f, 1.0f - dd) > 0.0f;
}

__global__ __launch_bounds__(256) void k_brick_required(ReqArgs A, uint8_t* __restrict__ flags) {
#pragma clang fp contract(off)
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int n = A.n;
  if (i >= int64_t(n) * n * n) return;
  const int x = int(i % n), y = int((i / n) % n), z = int(i / (int64_t(n) * n));
  bool need = emitted(A, x, y, z);
  const float px = float(x) + 0.5f, py = float(y) + 0.5f, pz = float(z) + 0.5f;
  const float mx = px - 0.5f * sample_face_cells(A, 0, px, py, pz);
  const float my = py - 0.5f * sample_face_cells(A, 1, px, py, pz);
  const float mz = pz - 0.5f * sample_face_cells(A, 2, px, py, pz);
  const float top = float(n - 1);
  const float qx = fminf(fmaxf(px - sample_face_cells(A, 0, mx, my, mz) - 0.5f, 0.0f), top);
  const float qy = fminf(fmaxf(py - sample_face_cells(A, 1, mx, my, mz) - 0.5f, 0.0f), top);
  const float qz = fminf(fmaxf(pz - sample_face_cells(A, 2, mx, my, mz) - 0.5f, 0.0f), top);
  const int bx = int(floorf(qx)), by = int(floorf(qy)), bz = int(floorf(qz));
  for (int dz = 0; dz < 2 && !need; ++dz)
    for (int dy = 0; dy < 2 && !need; ++dy)
      for (int dx = 0; dx < 2 && !need; ++dx) {
        const int cx = min(max(bx + dx, 0), n - 1), cy = min(max(by + dy, 0), n - 1), cz = min(max(bz + dz, 0), n - 1);
        need = A.density[(int64_t(cz) * n + cy) * n + cx] > 0.0f || emitted(A, cx, cy, cz);
      }
  if (need) flags[(int64_t(z / A.side) * A.grid + y / A.side) * A.grid + x / A.side] = 1;
}

// Brick map of a cubic n^3 domain: map[(bz*g + by)*g + bx] = slot or -1.
__global__ void k_brick_map(int n_active, const uint32_t* __restrict__ active, const int4* __restrict__ coords,
                            int g, int32_t* __restrict__ map) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_active) return;
  const uint32_t slot = active[t];
  const int4 c = coords[slot];
  if (c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x < g && c.y < g && c.z < g)
    map[(int64_t(c.z) * g + c.y) * g + c.x] = int32_t(slot);
}

// Semi-Lagrangian transport of the brick field (native/projection.mm:139-147
// transportDensity): every voxel of an active brick inside the n^3 domain is
// traced back through the MAC velocity (RK2, cells per step) and samples the
// old field trilinearly at the clamped donor cells; donors in inactive bricks
// read 0, voxels outside the domain get 0.  Same fp32 operations as the
// preflight (contraction off), so it equals the dense float32 transport
// wherever the active bricks cover the preflight's flags.
__global__ __launch_bounds__(256) void k_brick_transport(ReqArgs A, const uint32_t* __restrict__ active,
                                                         const int4* __restrict__ coords,
                                                         const int32_t* __restrict__ map,
                                                         const float* __restrict__ src, float* __restrict__ dst) {
#pragma clang fp contract(off)
  const uint32_t slot = active[blockIdx.x];
  const int4 bc = coords[slot];
  const int side = A.side, n = A.n, g = A.grid, vox = side * side * side;
  for (int l = threadIdx.x; l < vox; l += blockDim.x) {
    const int x = bc.x * side + l % side, y = bc.y * side + (l / side) % side, z = bc.z * side + l / (side * side);
    float value = 0.0f;
    if (x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n) {
      const float px = float(x) + 0.5f, py = float(y) + 0.5f, pz = float(z) + 0.5f;
      const float mx = px - 0.5f * sample_face_cells(A, 0, px, py, pz);
      const float my = py - 0.5f * sample_face_cells(A, 1, px, py, pz);
      const float mz = pz - 0.5f * sample_face_cells(A, 2, px, py, pz);
      const float top = float(n - 1);
      const float qx = fminf(fmaxf(px - sample_face_cells(A, 0, mx, my, mz) - 0.5f, 0.0f), top);
      const float qy = fminf(fmaxf(py - sample_face_cells(A, 1, mx, my, mz) - 0.5f, 0.0f), top);
      const float qz = fminf(fmaxf(pz - sample_face_cells(A, 2, mx, my, mz) - 0.5f, 0.0f), top);
      const float fbx = floorf(qx), fby = floorf(qy), fbz = floorf(qz);
      const float fx = qx - fbx, fy = qy - fby, fz = qz - fbz;
      const int bx = int(fbx), by = int(fby), bz = int(fbz);
      for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
          for (int dx = 0; dx < 2; ++dx) {
            const int cx = min(bx + dx, n - 1), cy = min(by + dy, n - 1), cz = min(bz + dz, n - 1);
            const int32_t s = map[(int64_t(cz / side) * g + cy / side) * g + cx / side];
            const float d = s < 0 ? 0.0f
                                  : src[size_t(s) * vox + ((cz % side) * side + cy % side) * side + cx % side];
            const float wx = dx ? fx : 1.0f - fx, wy = dy ? fy : 1.0f - fy, wz = dz ? fz : 1.0f - fz;
            const float t = wx * wy;
            const float wgt = t * wz;
            const float term = wgt * d;
            value = value + term;
          }
    }
    dst[size_t(slot) * vox + l] = value;
  }
}

// Diagnostic probe: every reserved voxel is set to the sentinel, then each
// active brick writes, per voxel, a pattern of its slot, voxel index and the
// six neighbour ids; the host verifies every slot.
__host__ __device__ inline uint32_t probe_pattern(uint32_t slot, uint32_t v, const int32_t* n6) {
  uint32_t h = slot * 0x9e3779b1u ^ (v + 0x7f4a7c15u);
  for (int d = 0; d < 6; ++d) h = (h ^ uint32_t(n6[d])) * 0x85ebca6bu + uint32_t(d);
  return h ^ (h >> 15);
}
constexpr uint32_t kProbeSentinel = 0xdeadbeefu;

__global__ __launch_bounds__(256) void k_brick_probe(int side, const uint32_t* __restrict__ active,
                                                     const int32_t* __restrict__ nbr, uint32_t* __restrict__ field) {
  const uint32_t slot = active[blockIdx.x];
  const int vox = side * side * side;
  int32_t n6[6];
  for (int d = 0; d < 6; ++d) n6[d] = nbr[size_t(slot) * 6 + d];
  for (int v = threadIdx.x; v < vox; v += blockDim.x) field[size_t(slot) * vox + v] = probe_pattern(slot, v, n6);
}

}  // namespace fx

using namespace fx;

// The pool: topology, regions and the device arrays.
struct fx_bricks {
  int32_t side = 8;
  uint64_t budget = 0;
  std::unique_ptr<BrickTopo> topo;
  std::unique_ptr<Regions> regions;
  // device
  float* field = nullptr;      // capacity * side^3
  int4* coords = nullptr;      // capacity
  uint32_t* active = nullptr;  // capacity
  int32_t* nbr = nullptr;      // capacity * 6
  uint64_t* keys = nullptr;    // table
  int32_t* vals = nullptr;     // table
  size_t table = 0;            // entries allocated on the device
  uint32_t dev_capacity = 0;
  uint32_t synced_active = 0;
  uint32_t reallocations = 0;
  float* field2 = nullptr;     // transport target, capacity * side^3 (on demand)
  uint32_t field2_capacity = 0;
  int32_t* map = nullptr;      // brick map of the last transport domain (on demand)
  int64_t map_cells = 0;
  uint32_t* fresh = nullptr;   // capacity: slots whose payload a sync clears
  bool closed = false;
};

namespace fx {

static uint64_t pool_bytes(uint32_t side, uint32_t capacity) {
  const uint64_t n = uint64_t(capacity) * (uint64_t(side) * side * side * 4 + 28);
  return (n + 255) & ~uint64_t(255);
}

static void release(fx_bricks* b) {
  for (void* p : {(void*)b->field, (void*)b->coords, (void*)b->active, (void*)b->nbr, (void*)b->keys,
                  (void*)b->vals, (void*)b->field2, (void*)b->map})
    if (p) (void)hipFree(p);
  b->field2 = nullptr;
  b->field2_capacity = 0;
  b->map = nullptr;
  b->map_cells = 0;
  b->field = nullptr;
  b->coords = nullptr;
  b->active = nullptr;
  b->nbr = nullptr;
  b->keys = nullptr;
  b->vals = nullptr;
  b->table = 0;
  if (b->fresh) (void)hipFree(b->fresh);
  b->fresh = nullptr;
  b->dev_capacity = 0;
}

// device arrays for the topology's capacity; the old field's values are kept
static int alloc_device(fx_bricks* b, hipStream_t s) {
  const uint32_t cap = b->topo->capacity();
  if (cap == b->dev_capacity && b->table == b->topo->keys().size()) return FX_OK;
  const size_t vox = size_t(b->side) * b->side * b->side;
  float* field = nullptr;
  FX_HIP(hipMalloc(&field, sizeof(float) * vox * cap));
  if (b->field) {
    FX_HIP(hipMemcpyAsync(field, b->field, sizeof(float) * vox * b->dev_capacity, hipMemcpyDeviceToDevice, s));
    FX_HIP(hipStreamSynchronize(s));
    ++b->reallocations;
  }
  float* old_field = b->field;
  b->field = nullptr;
  release(b);
  if (old_field) (void)hipFree(old_field);
  b->field = field;
  FX_HIP(hipMalloc(&b->coords, sizeof(int4) * cap));
  FX_HIP(hipMalloc(&b->active, sizeof(uint32_t) * cap));
  FX_HIP(hipMalloc(&b->nbr, sizeof(int32_t) * 6 * cap));
  FX_HIP(hipMalloc(&b->fresh, sizeof(uint32_t) * cap));
  b->table = b->topo->keys().size();
  FX_HIP(hipMalloc(&b->keys, sizeof(uint64_t) * b->table));
  FX_HIP(hipMalloc(&b->vals, sizeof(int32_t) * b->table));
  b->dev_capacity = cap;
  return FX_OK;
}

// one float4 per thread: 128 threads cover an 8^3 brick, 256 a quarter of 16^3
static unsigned copy4_threads(int side) { return side == 8 ? 128u : 256u; }

template <class F>
static int guarded(F&& f) {
  try {
    return f();
  } catch (const std::invalid_argument& e) {
    set_error("%s", e.what());
    return FX_EINVAL;
  } catch (const std::out_of_range& e) {
    set_error("%s", e.what());
    return FX_EINVAL;
  } catch (const std::bad_alloc&) {
    set_error("out of host memory");
    return FX_ESTATE;
  } catch (const std::exception& e) {
    set_error("%s", e.what());
    return FX_ESTATE;
  }
}

static int ready(const fx_bricks* b) {
  FX_REQUIRE(b, FX_EINVAL, "null brick pool");
  FX_REQUIRE(!b->closed, FX_ESTATE, "Brick pool is closed");
  return FX_OK;
}

}  // namespace fx

extern "C" {

int fx_bricks_create(int32_t side, uint32_t capacity, uint64_t budget_bytes, fx_bricks** out) {
  FX_REQUIRE(out, FX_EINVAL, "null output");
  *out = nullptr;
  FX_REQUIRE(side == 8 || side == 16, FX_EINVAL, "side: 8 or 16; capacity: 1..262144");
  FX_REQUIRE(capacity >= 1 && capacity <= kMaxBricks, FX_EINVAL, "side: 8 or 16; capacity: 1..262144");
  FX_REQUIRE(pool_bytes(uint32_t(side), capacity) <= budget_bytes, FX_EINVAL,
             "Brick payload and GPU topology exceed budget");
  return guarded([&]() -> int {
    auto b = std::make_unique<fx_bricks>();
    b->side = side;
    b->budget = budget_bytes;
    b->topo = std::make_unique<BrickTopo>(capacity);
    b->regions = std::make_unique<Regions>(capacity);
    *out = b.release();
    return FX_OK;
  });
}

int fx_bricks_destroy(fx_bricks* b) {
  if (!b) return FX_OK;
  release(b);
  delete b;
  return FX_OK;
}

int fx_bricks_close(fx_bricks* b) {
  if (int e = ready(b)) return e;
  release(b);
  b->closed = true;
  return FX_OK;
}

int fx_bricks_activate(fx_bricks* b, int32_t x, int32_t y, int32_t z, uint32_t* slot) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(slot, FX_EINVAL, "null output");
  return guarded([&]() -> int {
    *slot = b->topo->activate({x, y, z});
    return FX_OK;
  });
}

int fx_bricks_deactivate(fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t* removed) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(removed, FX_EINVAL, "null output");
  return guarded([&]() -> int {
    *removed = b->topo->deactivate({x, y, z}) ? 1 : 0;
    return FX_OK;
  });
}

int fx_bricks_lookup(const fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t* slot) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(slot, FX_EINVAL, "null output");
  return guarded([&]() -> int {
    *slot = b->topo->lookup({x, y, z});
    return FX_OK;
  });
}

int fx_bricks_neighbors(const fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t* out6) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(out6, FX_EINVAL, "null output");
  return guarded([&]() -> int {
    b->topo->neighbors({x, y, z}, out6);
    return FX_OK;
  });
}

int fx_bricks_activate_box(fx_bricks* b, int32_t x, int32_t y, int32_t z, int32_t nx, int32_t ny, int32_t nz) {
  if (int e = ready(b)) return e;
  return guarded([&]() -> int {
    b->topo->activate_box({x, y, z}, {nx, ny, nz});
    return FX_OK;
  });
}

int fx_bricks_activate_flags(fx_bricks* b, const uint8_t* flags, int32_t gx, int32_t gy, int32_t gz) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(flags && gx >= 1 && gy >= 1 && gz >= 1 && gx <= 4096 && gy <= 4096 && gz <= 4096, FX_EINVAL,
             "flag grid: 1..4096 bricks per axis");
  return guarded([&]() -> int {
    uint32_t needed = 0;
    for (int z = 0; z < gz; ++z)
      for (int y = 0; y < gy; ++y)
        for (int x = 0; x < gx; ++x)
          if (flags[(size_t(z) * gy + y) * gx + x]) needed += b->topo->lookup({x, y, z}) < 0;
    if (needed > b->topo->free_count()) throw std::runtime_error("Brick pool capacity exhausted");
    for (int z = 0; z < gz; ++z)
      for (int y = 0; y < gy; ++y)
        for (int x = 0; x < gx; ++x)
          if (flags[(size_t(z) * gy + y) * gx + x]) b->topo->activate({x, y, z});
    return FX_OK;
  });
}

int fx_bricks_update_regions(fx_bricks* b, const double* sources, int32_t count, int32_t grid, double dt,
                             int32_t linger, int32_t halo) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(count >= 0 && count <= 64, FX_EINVAL, "At most 64 sources");
  FX_REQUIRE(count == 0 || sources, FX_EINVAL, "null sources");
  return guarded([&]() -> int {
    std::vector<RegionSrc> src(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) {
      const double* r = sources + 7 * i;
      src[i] = RegionSrc{r[0], r[1], r[2], r[3], r[4], r[5], r[6]};
    }
    b->regions->update(*b->topo, src, grid, dt, linger, halo);
    return FX_OK;
  });
}

int fx_bricks_count(const fx_bricks* b, uint32_t* active, uint32_t* capacity) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(active && capacity, FX_EINVAL, "null output");
  *active = uint32_t(b->topo->active().size());
  *capacity = b->topo->capacity();
  return FX_OK;
}

int fx_bricks_snapshot(const fx_bricks* b, int32_t* rows, uint32_t max_rows, uint32_t* count) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(count, FX_EINVAL, "null output");
  const auto& act = b->topo->active();
  *count = uint32_t(act.size());
  FX_REQUIRE(max_rows >= act.size() && (act.empty() || rows), FX_EINVAL, "snapshot needs %u rows",
             unsigned(act.size()));
  return guarded([&]() -> int {
    for (size_t i = 0; i < act.size(); ++i) {
      const BC c = b->topo->coord(act[i]);
      int32_t* r = rows + 5 * i;
      r[0] = c.x;
      r[1] = c.y;
      r[2] = c.z;
      r[3] = int32_t(act[i]);
      r[4] = int32_t(b->regions->age(act[i]));
    }
    return FX_OK;
  });
}

int fx_bricks_host_bytes(const fx_bricks* b, uint64_t* bytes) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(bytes, FX_EINVAL, "null output");
  *bytes = b->topo->metadata_bytes() + b->regions->bytes();
  return FX_OK;
}

int fx_bricks_reserve(fx_bricks* b, uint32_t needed, uint32_t limit, uint32_t* capacity) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(capacity, FX_EINVAL, "null output");
  FX_REQUIRE(limit >= 1 && limit <= kMaxBricks, FX_EINVAL, "capacity limit must be 1..262144");
  FX_REQUIRE(needed <= limit, FX_ESTATE, "Brick pool capacity exhausted");
  uint32_t cap = b->topo->capacity();
  while (cap < needed) cap = uint32_t(std::min<uint64_t>(uint64_t(cap) * 2, limit));
  FX_REQUIRE(pool_bytes(uint32_t(b->side), cap) <= b->budget, FX_ESTATE,
             "Brick capacity %u exceeds the pool budget", unsigned(cap));
  if (cap > b->topo->capacity()) {
    b->topo->grow(cap);
    b->regions->grow(cap);
  }
  *capacity = cap;
  return FX_OK;
}

int fx_bricks_sync(fx_bricks* b, void* stream) {
  if (int e = ready(b)) return e;
  hipStream_t s = S(stream);
  if (int e = alloc_device(b, s)) return e;
  const BrickTopo& t = *b->topo;
  const uint32_t cap = t.capacity();
  std::vector<int4> coords(cap);
  for (uint32_t i = 0; i < cap; ++i) {
    const BC c = t.coord(i);
    coords[i] = make_int4(c.x, c.y, c.z, 0);
  }
  const auto& act = t.active();
  FX_HIP(hipMemcpyAsync(b->coords, coords.data(), sizeof(int4) * cap, hipMemcpyHostToDevice, s));
  if (!act.empty())
    FX_HIP(hipMemcpyAsync(b->active, act.data(), sizeof(uint32_t) * act.size(), hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(b->keys, t.keys().data(), sizeof(uint64_t) * b->table, hipMemcpyHostToDevice, s));
  FX_HIP(hipMemcpyAsync(b->vals, t.vals().data(), sizeof(int32_t) * b->table, hipMemcpyHostToDevice, s));
  // newly activated slots start empty, as new brick storage does in the reference
  std::vector<uint32_t> fresh;
  {
    std::vector<char> seen(cap, 0);
    for (uint32_t slot : b->topo->take_fresh())
      if (slot < cap && !seen[slot]) {
        seen[slot] = 1;
        fresh.push_back(slot);
      }
  }
  if (!fresh.empty()) {
    FX_HIP(hipMemcpyAsync(b->fresh, fresh.data(), sizeof(uint32_t) * fresh.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_brick_clear, dim3(unsigned(fresh.size())), dim3(256), 0, s, b->side * b->side * b->side,
                       b->fresh, b->field);
    FX_LAUNCHED("k_brick_clear");
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, s, int64_t(cap) * 6, b->nbr, -1);
  const int n = int(act.size());
  if (n) {
    hipLaunchKernelGGL(k_brick_neighbors, dim3((n * 6 + 255) / 256), dim3(256), 0, s, n, b->active, b->coords,
                       b->keys, b->vals, uint32_t(b->table - 1), b->nbr);
  }
  FX_LAUNCHED("k_brick_neighbors");
  // the host vectors above are pageable: finish the copies before they go
  FX_HIP(hipStreamSynchronize(s));
  b->synced_active = uint32_t(n);
  return FX_OK;
}

int fx_bricks_device(const fx_bricks* b, fx_bricks_dev* out) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(out, FX_EINVAL, "null output");
  FX_REQUIRE(b->field, FX_ESTATE, "brick pool has no device arrays: call fx_bricks_sync first");
  out->field = b->field;
  out->coords = reinterpret_cast<int32_t*>(b->coords);
  out->active = b->active;
  out->neighbors = b->nbr;
  out->side = b->side;
  out->capacity = b->dev_capacity;
  out->active_count = b->synced_active;
  out->reallocations = b->reallocations;
  return FX_OK;
}

int fx_bricks_gather(fx_bricks* b, int32_t nx, int32_t ny, int32_t nz, const float* dense, void* stream) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(dense && nx >= 1 && ny >= 1 && nz >= 1, FX_EINVAL, "gather: bad dense grid");
  FX_REQUIRE(b->field, FX_ESTATE, "brick pool has no device arrays: call fx_bricks_sync first");
  if (!b->synced_active) return FX_OK;
  if (nx % 4 == 0 && reinterpret_cast<uintptr_t>(dense) % 16 == 0)
    hipLaunchKernelGGL((k_brick_copy4<true>), dim3(b->synced_active), dim3(copy4_threads(b->side)), 0, S(stream), b->side, b->active,
                       b->coords, nx, ny, nz, reinterpret_cast<float4*>(const_cast<float*>(dense)),
                       reinterpret_cast<float4*>(b->field));
  else
    hipLaunchKernelGGL((k_brick_copy<true>), dim3(b->synced_active), dim3(256), 0, S(stream), b->side, b->active,
                       b->coords, nx, ny, nz, const_cast<float*>(dense), b->field);
  FX_LAUNCHED("k_brick_copy gather");
  return FX_OK;
}

int fx_bricks_scatter(fx_bricks* b, int32_t nx, int32_t ny, int32_t nz, float* dense, void* stream) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(dense && nx >= 1 && ny >= 1 && nz >= 1, FX_EINVAL, "scatter: bad dense grid");
  FX_REQUIRE(b->field, FX_ESTATE, "brick pool has no device arrays: call fx_bricks_sync first");
  if (!b->synced_active) return FX_OK;
  if (nx % 4 == 0 && reinterpret_cast<uintptr_t>(dense) % 16 == 0)
    hipLaunchKernelGGL((k_brick_copy4<false>), dim3(b->synced_active), dim3(copy4_threads(b->side)), 0, S(stream), b->side, b->active,
                       b->coords, nx, ny, nz, reinterpret_cast<float4*>(dense), reinterpret_cast<float4*>(b->field));
  else
    hipLaunchKernelGGL((k_brick_copy<false>), dim3(b->synced_active), dim3(256), 0, S(stream), b->side, b->active,
                       b->coords, nx, ny, nz, dense, b->field);
  FX_LAUNCHED("k_brick_copy scatter");
  return FX_OK;
}

int fx_brick_support(int32_t side, int32_t nx, int32_t ny, int32_t nz, const float* dense, float threshold,
                     uint8_t* flags, void* stream) {
  FX_REQUIRE(side == 8 || side == 16, FX_EINVAL, "side: 8 or 16");
  FX_REQUIRE(dense && flags && nx >= 1 && ny >= 1 && nz >= 1, FX_EINVAL, "support: bad arrays");
  FX_REQUIRE(threshold >= 0.0f, FX_EINVAL, "support: threshold must be >= 0");
  const int bx = (nx + side - 1) / side, by = (ny + side - 1) / side, bz = (nz + side - 1) / side;
  FX_REQUIRE(int64_t(bx) * by * bz < (int64_t(1) << 31), FX_EINVAL, "support: grid too large");
  if (nx % 4 == 0 && reinterpret_cast<uintptr_t>(dense) % 16 == 0) {
    const int64_t nq = int64_t(nx / 4) * ny * nz;
    hipLaunchKernelGGL(k_fill_u8, dim3(64), dim3(256), 0, S(stream), int64_t(bx) * by * bz, flags);
    const int64_t blocks = std::min<int64_t>((nq + 255) / 256, 8192);
    hipLaunchKernelGGL(k_brick_support_rows, dim3(unsigned(blocks)), dim3(256), 0, S(stream), side, nx, ny, bx, by, nq,
                       reinterpret_cast<const float4*>(dense), threshold, flags);
  } else {
    hipLaunchKernelGGL(k_brick_support, dim3(bx * by * bz), dim3(256), 0, S(stream), side, nx, ny, nz, bx, by, bz,
                       dense, threshold, flags);
  }
  FX_LAUNCHED("k_brick_support");
  return FX_OK;
}

int fx_brick_required(int32_t side, int32_t n, const float* density, const float* u, const float* v, const float* w,
                      const float* emitter, uint8_t* flags, void* stream) {
  FX_REQUIRE(side == 8 || side == 16, FX_EINVAL, "side: 8 or 16");
  FX_REQUIRE(n >= 1 && n <= 1024 && n % side == 0, FX_EINVAL, "required: n must be 1..1024 and a multiple of side");
  FX_REQUIRE(density && u && v && w && emitter && flags, FX_EINVAL, "required: null array");
  ReqArgs A{n, side, n / side, density, u, v, w, emitter[0], emitter[1], emitter[2], emitter[3], emitter[4]};
  FX_REQUIRE(A.radius > 0.0f && std::isfinite(A.cx) && std::isfinite(A.cy) && std::isfinite(A.cz) &&
                 std::isfinite(A.rate),
             FX_EINVAL, "required: emitter (cx, cy, cz, radius > 0, rate) in cells");
  const int64_t cells = int64_t(n) * n * n;
  hipLaunchKernelGGL(k_fill_u8, dim3(64), dim3(256), 0, S(stream), int64_t(A.grid) * A.grid * A.grid, flags);
  hipLaunchKernelGGL(k_brick_required, dim3(unsigned((cells + 255) / 256)), dim3(256), 0, S(stream), A, flags);
  FX_LAUNCHED("k_brick_required");
  return FX_OK;
}

int fx_bricks_transport(fx_bricks* b, int32_t n, const float* u, const float* v, const float* w, void* stream) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(n >= 1 && n <= 1024 && n % b->side == 0, FX_EINVAL, "transport: n must be 1..1024 and a multiple of side");
  FX_REQUIRE(u && v && w, FX_EINVAL, "transport: null velocity");
  FX_REQUIRE(b->field, FX_ESTATE, "brick pool has no device arrays: call fx_bricks_sync first");
  hipStream_t s = S(stream);
  const size_t vox = size_t(b->side) * b->side * b->side;
  if (b->field2_capacity != b->dev_capacity) {
    if (b->field2) FX_HIP(hipFree(b->field2));
    b->field2 = nullptr;
    b->field2_capacity = 0;
    FX_HIP(hipMalloc(&b->field2, sizeof(float) * vox * b->dev_capacity));
    b->field2_capacity = b->dev_capacity;
  }
  const int g = n / b->side;
  const int64_t cells = int64_t(g) * g * g;
  if (b->map_cells < cells) {
    if (b->map) FX_HIP(hipFree(b->map));
    b->map = nullptr;
    b->map_cells = 0;
    FX_HIP(hipMalloc(&b->map, sizeof(int32_t) * cells));
    b->map_cells = cells;
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(256), dim3(256), 0, s, cells, b->map, -1);
  if (b->synced_active) {
    hipLaunchKernelGGL(k_brick_map, dim3((b->synced_active + 255) / 256), dim3(256), 0, s, int(b->synced_active),
                       b->active, b->coords, g, b->map);
    ReqArgs A{n, b->side, g, nullptr, u, v, w, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
    hipLaunchKernelGGL(k_brick_transport, dim3(b->synced_active), dim3(256), 0, s, A, b->active,
                       b->coords, b->map, b->field, b->field2);
  }
  FX_LAUNCHED("k_brick_transport");
  std::swap(b->field, b->field2);
  return FX_OK;
}

int fx_bricks_probe(fx_bricks* b, uint32_t* mismatches, float* gpu_ms, void* stream) {
  if (int e = ready(b)) return e;
  FX_REQUIRE(mismatches && gpu_ms, FX_EINVAL, "null output");
  FX_REQUIRE(b->field, FX_ESTATE, "brick pool has no device arrays: call fx_bricks_sync first");
  hipStream_t s = S(stream);
  const size_t vox = size_t(b->side) * b->side * b->side, total = vox * b->dev_capacity;
  uint32_t* field = reinterpret_cast<uint32_t*>(b->field);
  hipLaunchKernelGGL(k_fill_i32, dim3(1024), dim3(256), 0, s, int64_t(total), reinterpret_cast<int32_t*>(field),
                     int32_t(kProbeSentinel));
  hipEvent_t e0, e1;
  FX_HIP(hipEventCreate(&e0));
  FX_HIP(hipEventCreate(&e1));
  FX_HIP(hipEventRecord(e0, s));
  if (b->synced_active)
    hipLaunchKernelGGL(k_brick_probe, dim3(b->synced_active), dim3(256), 0, s, b->side, b->active, b->nbr, field);
  FX_HIP(hipEventRecord(e1, s));
  FX_LAUNCHED("k_brick_probe");
  std::vector<uint32_t> host(total);
  std::vector<int32_t> nbr(size_t(b->dev_capacity) * 6);
  FX_HIP(hipMemcpyAsync(host.data(), field, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, s));
  FX_HIP(hipMemcpyAsync(nbr.data(), b->nbr, sizeof(int32_t) * nbr.size(), hipMemcpyDeviceToHost, s));
  FX_HIP(hipStreamSynchronize(s));
  FX_HIP(hipEventElapsedTime(gpu_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // expected: the host topology's neighbours for active slots, sentinels elsewhere
  std::vector<uint8_t> live(b->dev_capacity, 0);
  uint32_t bad = 0;
  for (uint32_t slot : b->topo->active()) {
    if (slot >= b->dev_capacity) continue;
    live[slot] = 1;
    int32_t n6[6];
    b->topo->neighbors(b->topo->coord(slot), n6);
    for (int d = 0; d < 6; ++d) bad += nbr[size_t(slot) * 6 + d] != n6[d];
    for (size_t v = 0; v < vox; ++v) bad += host[slot * vox + v] != probe_pattern(slot, uint32_t(v), n6);
  }
  for (uint32_t slot = 0; slot < b->dev_capacity; ++slot)
    if (!live[slot])
      for (size_t v = 0; v < vox; ++v) bad += host[slot * vox + v] != kProbeSentinel;
  *mismatches = bad;
  return FX_OK;
}

}  // extern "C"
