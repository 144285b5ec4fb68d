// Addressing of fields stored on cubic bricks, shared by the brick
// diagnostics (projection, coupled step, scalar and MAC advection).
//
// A level of n^3 unit cells is stored either densely (x-fast, one n^3
// "brick", sparse = 0) or as `side`^3 bricks in slots: cell (x, y, z) of the
// brick in slot s lives at s*side^3 + (lz*side + ly)*side + lx, with a
// grid^3 map from brick coordinate to slot (-1: missing) and a slot -> brick
// coordinate table.  Face arrays of axis a store (side+1)*side^2 faces per
// brick (the block is one face longer along a) at a*faces + s*(side+1)*side^2
// + (lz*dy + ly)*dx + lx.  The face at local index `side` along a repeats the
// next brick's face 0 and belongs to that brick (canonical ownership); only
// the last brick along a owns its upper face, the others leave it as padding
// that nothing reads or writes (native/mac.mm:21-27, projection.mm:35-41).
#pragma once
#include <vector>

#include "common.hpp"

namespace fx {
namespace brick {

struct Level {
  int n, side, grid, sparse;
  int cells, faces;   // stored cells; stored faces per axis
  int mode;           // solve region (see solve())
  const int* map;     // grid^3 brick -> slot (-1: missing)
  const int* coords;  // slot -> brick (x, y, z)
};

__host__ __device__ inline bool inside(int x, int y, int z, int n) {
  return x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n;
}

// cell of the solve region of a mode (projection.mm:21-27): 0, 3, 4 the
// whole box; 1 the central half [n/4, 3n/4)^3; 2 the boxes [n/8, 3n/8)^3 and
// [5n/8, 7n/8)^3
__host__ __device__ inline bool solve(int x, int y, int z, int n, int mode) {
  if (!inside(x, y, z, n)) return false;
  if (mode != 1 && mode != 2) return true;
  auto box = [&](int lo, int hi) { return x >= lo && y >= lo && z >= lo && x < hi && y < hi && z < hi; };
  if (mode == 1) return box(n / 4, 3 * n / 4);
  const int e = n / 4;
  return box(n / 8, n / 8 + e) || box(5 * n / 8, 5 * n / 8 + e);
}

__device__ __forceinline__ int thread_id() { return int(blockIdx.x * blockDim.x + threadIdx.x); }

__device__ __forceinline__ int cell_index(const Level& L, int x, int y, int z) {
  if (!inside(x, y, z, L.n)) return -1;
  if (!L.sparse) return (z * L.n + y) * L.n + x;
  const int s = L.side, bx = x / s, by = y / s, bz = z / s;
  const int slot = L.map[(bz * L.grid + by) * L.grid + bx];
  if (slot < 0) return -1;
  return slot * s * s * s + ((z - bz * s) * s + (y - by * s)) * s + (x - bx * s);
}

// coordinates of stored cell i
__device__ __forceinline__ void cell_coord(const Level& L, int i, int& x, int& y, int& z) {
  const int s = L.sparse ? L.side : L.n, per = s * s * s;
  const int slot = i / per, j = i - slot * per;
  x = j % s;
  y = (j / s) % s;
  z = j / (s * s);
  if (L.sparse) {
    x += L.coords[3 * slot] * s;
    y += L.coords[3 * slot + 1] * s;
    z += L.coords[3 * slot + 2] * s;
  }
}

// stored index of face (x, y, z) of axis a, or -1 (out of range / missing brick)
__device__ __forceinline__ int face_index(const Level& L, int a, int x, int y, int z) {
  int c[3] = {x, y, z};
  for (int k = 0; k < 3; ++k)
    if (c[k] < 0 || c[k] >= L.n + (k == a)) return -1;
  if (!L.sparse) {
    const int dx = L.n + (a == 0), dy = L.n + (a == 1);
    return a * L.faces + (z * dy + y) * dx + x;
  }
  const int s = L.side;
  int b[3] = {x / s, y / s, z / s};
  b[a] = min(b[a], L.grid - 1);  // the last face along a belongs to the last brick
  const int slot = L.map[(b[2] * L.grid + b[1]) * L.grid + b[0]];
  if (slot < 0) return -1;
  const int lx = x - b[0] * s, ly = y - b[1] * s, lz = z - b[2] * s;
  const int dx = s + (a == 0), dy = s + (a == 1);
  return a * L.faces + slot * (s + 1) * s * s + (lz * dy + ly) * dx + lx;
}

// global face coordinates of stored face i of axis a; false for padding
__device__ __forceinline__ bool face_coord(const Level& L, int a, int i, int c[3]) {
  const int s = L.sparse ? L.side : L.n, per = (s + 1) * s * s;
  const int slot = i / per, j = i - slot * per;
  const int dx = s + (a == 0), dy = s + (a == 1);
  c[0] = j % dx;
  c[1] = (j / dx) % dy;
  c[2] = j / (dx * dy);
  if (!L.sparse) return true;
  int b[3];
  for (int k = 0; k < 3; ++k) b[k] = L.coords[3 * slot + k];
  if (c[a] == s && b[a] != L.grid - 1) return false;
  for (int k = 0; k < 3; ++k) c[k] += b[k] * s;
  return true;
}

// cell value with q = 0 off the solve region and on missing bricks
__device__ __forceinline__ float cell_at(const Level& L, const float* q, int x, int y, int z) {
  if (!solve(x, y, z, L.n, L.mode)) return 0.0f;
  const int i = cell_index(L, x, y, z);
  return i < 0 ? 0.0f : q[i];
}

__device__ __forceinline__ float face_at(const Level& L, const float* u, int a, int x, int y, int z) {
  const int i = face_index(L, a, x, y, z);
  return i < 0 ? 0.0f : u[i];
}

__device__ __forceinline__ float corner_weight(int d, float f) { return d ? f : 1.0f - f; }

// trilinear sample of face array a at face-index coordinates q: q clamped to
// [0, size-1], corner index min(b+1, size-1) (mac.mm:29-35)
__device__ __forceinline__ float sample_face(const Level& L, const float* u, int a, float qx, float qy, float qz) {
  const int sx = L.n + (a == 0), sy = L.n + (a == 1), sz = L.n + (a == 2);
  qx = fminf(fmaxf(qx, 0.0f), float(sx - 1));
  qy = fminf(fmaxf(qy, 0.0f), float(sy - 1));
  qz = fminf(fmaxf(qz, 0.0f), float(sz - 1));
  const int bx = int(floorf(qx)), by = int(floorf(qy)), bz = int(floorf(qz));
  const float fx = qx - bx, fy = qy - by, fz = qz - bz;
  float v = 0.0f;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const float w = corner_weight(dx, fx) * corner_weight(dy, fy) * corner_weight(dz, fz);
        v = __builtin_fmaf(w, face_at(L, u, a, min(bx + dx, sx - 1), min(by + dy, sy - 1), min(bz + dz, sz - 1)), v);
      }
  return v;
}

// staggered velocity at a point in cell units (mac.mm:36-39)
__device__ __forceinline__ void velocity_at(const Level& L, const float* u, const float p[3], float v[3]) {
  v[0] = sample_face(L, u, 0, p[0], p[1] - 0.5f, p[2] - 0.5f);
  v[1] = sample_face(L, u, 1, p[0] - 0.5f, p[1], p[2] - 0.5f);
  v[2] = sample_face(L, u, 2, p[0] - 0.5f, p[1] - 0.5f, p[2]);
}

// RK2 midpoint back-trace with dt = 1 cell: p - v(p - v(p)/2)
__device__ __forceinline__ void departure(const Level& L, const float* u, const float p[3], float d[3]) {
  float v[3], mid[3];
  velocity_at(L, u, p, v);
  for (int k = 0; k < 3; ++k) mid[k] = __builtin_fmaf(-0.5f, v[k], p[k]);
  velocity_at(L, u, mid, v);
  for (int k = 0; k < 3; ++k) d[k] = p[k] - v[k];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

struct DeviceBuffer {
  void* p = nullptr;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};

// bricks of the box [lo, hi)^3 (brick units) in x-fast order
inline void box_topology(int grid, const int lo[3], const int hi[3], std::vector<int>& map, std::vector<int>& coords) {
  map.assign(size_t(grid) * grid * grid, -1);
  coords.clear();
  for (int z = lo[2]; z < hi[2]; ++z)
    for (int y = lo[1]; y < hi[1]; ++y)
      for (int x = lo[0]; x < hi[0]; ++x) {
        map[(size_t(z) * grid + y) * grid + x] = int(coords.size() / 3);
        coords.insert(coords.end(), {x, y, z});
      }
}

// dense x-fast index of face (x, y, z) of axis a on an n^3 grid
inline size_t dense_face(int n, int a, int x, int y, int z) {
  const size_t dx = n + (a == 0), dy = n + (a == 1);
  return (size_t(z) * dy + y) * dx + x;
}

}  // namespace brick
}  // namespace fx
