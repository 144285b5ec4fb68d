// Unit-cell impulse projection on 8^3 / 16^3 brick storage with canonical MAC
// face ownership, next to the same projection on dense storage (the brick
// engine's diagnostic comparison, native/projection.mm:9-168, 169-462).
//
// Storage of a level with `side`-cell bricks: cell (x, y, z) of the brick in
// slot s lives at s*side^3 + (lz*side + ly)*side + lx; face (x, y, z) of axis
// a at a*F + s*(side+1)*side^2 + local index in the (side+1 along a) block.
// A brick stores both faces of its cells along every axis, so divergence
// stays inside one brick; the face at local index `side` along a repeats the
// next brick's face 0 and is owned by that brick (canonical ownership) unless
// the brick is the last along a: such padding faces are never written.
// Dense storage is the same code with one n^3 "brick" (sparse = 0).
//
// The solve region is analytic (a mode), the pressure is 0 outside it and
// on missing bricks, the outer walls are Neumann, and velocity faces outside
// the bricks read 0.  Jacobi and the V-cycle follow the reference engine:
// omega = 2/3, 4 + 4 sweeps per level, residual restricted as 4 x the
// 8-child mean (0.5 x the sum), cell-centred trilinear prolongation with
// clamped indices, 128 sweeps on the 8^3 level.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <vector>

#include "common.hpp"

namespace fx {
namespace bp {

struct Level {
  int n, side, grid, sparse;
  int cells, faces;  // stored cells; stored faces per axis
  int mode;
  const int* map;     // grid^3 brick -> slot (-1: missing)
  const int* coords;  // slot -> brick (x, y, z)
};

__host__ __device__ inline bool inside(int x, int y, int z, int n) {
  return x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n;
}

// cell of the solve region (projection.mm:21-27)
__host__ __device__ inline bool solve(int x, int y, int z, int n, int mode) {
  if (!inside(x, y, z, n)) return false;
  if (mode != 1 && mode != 2) return true;
  auto box = [&](int lo, int hi) { return x >= lo && y >= lo && z >= lo && x < hi && y < hi && z < hi; };
  if (mode == 1) return box(n / 4, 3 * n / 4);
  const int e = n / 4;
  return box(n / 8, n / 8 + e) || box(5 * n / 8, 5 * n / 8 + e);
}

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

__device__ __forceinline__ float pressure_at(const Level& L, const float* q, int x, int y, int z) {
  if (!solve(x, y, z, L.n, L.mode)) return 0.0f;
  const int i = cell_index(L, x, y, z);
  return i < 0 ? 0.0f : q[i];
}

__device__ __forceinline__ float face_at(const Level& L, const float* u, int a, int x, int y, int z) {
  const int i = face_index(L, a, x, y, z);
  return i < 0 ? 0.0f : u[i];
}

__device__ __forceinline__ int thread_id() { return int(blockIdx.x * blockDim.x + threadIdx.x); }

// sum_a (u_a[c + e_a] - u_a[c]) on solve cells
__global__ void k_divergence(Level L, const float* __restrict__ u, float* __restrict__ out) {
  const int i = thread_id();
  if (i >= L.cells) return;
  int x, y, z;
  cell_coord(L, i, x, y, z);
  float v = 0.0f;
  if (solve(x, y, z, L.n, L.mode)) {
    v += face_at(L, u, 0, x + 1, y, z) - face_at(L, u, 0, x, y, z);
    v += face_at(L, u, 1, x, y + 1, z) - face_at(L, u, 1, x, y, z);
    v += face_at(L, u, 2, x, y, z + 1) - face_at(L, u, 2, x, y, z);
  }
  out[i] = v;
}

// one damped-Jacobi sweep of L q = rhs: in-box neighbours count in the
// diagonal, q = 0 outside the solve region
__global__ void k_jacobi(Level L, const float* __restrict__ q, const float* __restrict__ rhs, float* __restrict__ out) {
  const int i = thread_id();
  if (i >= L.cells) return;
  int x, y, z;
  cell_coord(L, i, x, y, z);
  if (!solve(x, y, z, L.n, L.mode)) {
    out[i] = 0.0f;
    return;
  }
  float sum = 0.0f, diag = 0.0f;
  const int c[3] = {x, y, z};
  for (int a = 0; a < 3; ++a)
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      int v[3] = {c[0], c[1], c[2]};
      v[a] += sgn;
      if (inside(v[0], v[1], v[2], L.n)) {
        diag += 1.0f;
        sum += pressure_at(L, q, v[0], v[1], v[2]);
      }
    }
  const float old = q[i];
  out[i] = __builtin_fmaf(2.0f / 3.0f, (sum - rhs[i]) / diag - old, old);
}

// rhs - L q on solve cells
__global__ void k_residual(Level L, const float* __restrict__ q, const float* __restrict__ rhs,
                           float* __restrict__ out) {
  const int i = thread_id();
  if (i >= L.cells) return;
  int x, y, z;
  cell_coord(L, i, x, y, z);
  if (!solve(x, y, z, L.n, L.mode)) {
    out[i] = 0.0f;
    return;
  }
  float lap = 0.0f;
  const int c[3] = {x, y, z};
  for (int a = 0; a < 3; ++a)
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      int v[3] = {c[0], c[1], c[2]};
      v[a] += sgn;
      if (inside(v[0], v[1], v[2], L.n)) lap += pressure_at(L, q, v[0], v[1], v[2]) - q[i];
    }
  out[i] = rhs[i] - lap;
}

// interior faces with a solve cell on either side: u - (q[c] - q[c - e_a]);
// other owned faces 0; padding faces untouched.  blockIdx.y = axis.
__global__ void k_gradient(Level L, const float* __restrict__ u, const float* __restrict__ q, float* __restrict__ out) {
  const int a = blockIdx.y;
  const int i = thread_id();
  if (i >= L.faces) return;
  const int s = L.sparse ? L.side : L.n, per = (s + 1) * s * s;
  const int slot = i / per, j = i - slot * per;
  const int dx = s + (a == 0), dy = s + (a == 1);
  int c[3] = {j % dx, (j / dx) % dy, j / (dx * dy)};
  int b[3] = {0, 0, 0};
  if (L.sparse) {
    for (int k = 0; k < 3; ++k) b[k] = L.coords[3 * slot + k];
    if (c[a] == s && b[a] != L.grid - 1) return;  // padding: owned by the next brick
  }
  for (int k = 0; k < 3; ++k) c[k] += b[k] * s;
  int lo[3] = {c[0], c[1], c[2]};
  lo[a] -= 1;
  float v = 0.0f;
  if (c[a] > 0 && c[a] < L.n && (solve(c[0], c[1], c[2], L.n, L.mode) || solve(lo[0], lo[1], lo[2], L.n, L.mode)))
    v = u[a * L.faces + i] - (pressure_at(L, q, c[0], c[1], c[2]) - pressure_at(L, q, lo[0], lo[1], lo[2]));
  out[a * L.faces + i] = v;
}

// coarse rhs = 0.5 * sum of the 8 children's residuals (4 x their mean: the
// coarse spacing is doubled) on coarse solve cells
__global__ void k_restrict(Level F, Level C, const float* __restrict__ fine, float* __restrict__ coarse) {
  const int i = thread_id();
  if (i >= C.cells) return;
  int x, y, z;
  cell_coord(C, i, x, y, z);
  float sum = 0.0f;
  if (solve(x, y, z, C.n, C.mode))
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) sum += pressure_at(F, fine, 2 * x + dx, 2 * y + dy, 2 * z + dz);
  coarse[i] = sum * 0.5f;
}

// fine q += cell-centred trilinear interpolation of the coarse correction
// (indices clamped to the coarse level) on fine solve cells
__global__ void k_prolong(Level F, Level C, const float* __restrict__ coarse, float* __restrict__ fine) {
  const int i = thread_id();
  if (i >= F.cells) return;
  int x, y, z;
  cell_coord(F, i, x, y, z);
  if (!solve(x, y, z, F.n, F.mode)) return;
  const float px = (x + 0.5f) * 0.5f - 0.5f, py = (y + 0.5f) * 0.5f - 0.5f, pz = (z + 0.5f) * 0.5f - 0.5f;
  const int bx = int(floorf(px)), by = int(floorf(py)), bz = int(floorf(pz));
  const float fx = px - bx, fy = py - by, fz = pz - bz;
  float sum = 0.0f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int qx = min(max(bx + dx, 0), C.n - 1), qy = min(max(by + dy, 0), C.n - 1),
                  qz = min(max(bz + dz, 0), C.n - 1);
        const float w = (dx ? fx : 1.0f - fx) * (dy ? fy : 1.0f - fy) * (dz ? fz : 1.0f - fz);
        sum = __builtin_fmaf(w, pressure_at(C, coarse, qx, qy, qz), sum);
      }
  fine[i] += sum;
}

// ---------------------------------------------------------------------------
// host engine
// ---------------------------------------------------------------------------
constexpr float kPadding = 12345.0f;  // value of never-written padding faces

static double exact_q(int x, int y, int z, int n) {
  const double k = 8.0 * M_PI / n;
  return std::cos(k * (x + 0.5)) * std::cos(k * (y + 0.5)) * std::cos(k * (z + 0.5));
}

// seeded face velocity of a mode (projection.mm:155-167)
static float seed_face(int x, int y, int z, int a, int n, int mode) {
  int c[3] = {x, y, z}, lo[3] = {x, y, z};
  lo[a] -= 1;
  if (c[a] == 0 || c[a] == n || mode == 3 ||
      (!solve(c[0], c[1], c[2], n, mode) && !solve(lo[0], lo[1], lo[2], n, mode)))
    return 0.0f;
  if (mode == 4) return float(exact_q(x, y, z, n)) - float(exact_q(lo[0], lo[1], lo[2], n));
  float pos[3] = {x + 0.5f, y + 0.5f, z + 0.5f};
  pos[a] -= 0.5f;
  return float(0.2 * (a + 1) * std::sin(8.0 * M_PI * pos[a] / n) + 0.1 * std::cos(4.0 * M_PI * pos[(a + 1) % 3] / n));
}

static uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

struct DeviceBuffer {
  void* p = nullptr;
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
};

struct Engine {
  Level L{};
  std::vector<int> map, coords;  // host copies
  uint64_t bytes = 0;
  DeviceBuffer mem;
  float *u_in = nullptr, *u_out = nullptr, *q[2] = {nullptr, nullptr}, *rhs = nullptr, *after = nullptr;
  int current = 0;

  Engine(int n, int side, int mode, bool sparse, std::vector<int> bmap, std::vector<int> bcoords, bool velocity) {
    L.n = n;
    L.side = sparse ? side : n;
    L.grid = sparse ? n / side : 1;
    L.sparse = sparse ? 1 : 0;
    L.mode = mode;
    const int bricks = sparse ? int(bcoords.size() / 3) : 1;
    const int s = L.side;
    L.cells = bricks * s * s * s;
    L.faces = velocity ? bricks * (s + 1) * s * s : 0;
    map = std::move(bmap);
    coords = std::move(bcoords);
    const uint64_t vel = align256(uint64_t(L.faces) * 3 * 4), scal = align256(uint64_t(L.cells) * 4);
    bytes = 2 * vel + 4 * scal + (sparse ? align256(map.size() * 4) + align256(coords.size() * 4) : 0);
  }

  int allocate(hipStream_t st) {
    FX_HIP(hipMalloc(&mem.p, bytes));
    char* b = static_cast<char*>(mem.p);
    const uint64_t vel = align256(uint64_t(L.faces) * 3 * 4), scal = align256(uint64_t(L.cells) * 4);
    u_in = reinterpret_cast<float*>(b);
    u_out = reinterpret_cast<float*>(b + vel);
    q[0] = reinterpret_cast<float*>(b + 2 * vel);
    q[1] = reinterpret_cast<float*>(b + 2 * vel + scal);
    rhs = reinterpret_cast<float*>(b + 2 * vel + 2 * scal);
    after = reinterpret_cast<float*>(b + 2 * vel + 3 * scal);
    FX_HIP(hipMemsetAsync(mem.p, 0, bytes, st));
    if (L.sparse) {
      int* m = reinterpret_cast<int*>(b + 2 * vel + 4 * scal);
      int* c = reinterpret_cast<int*>(b + 2 * vel + 4 * scal + align256(map.size() * 4));
      FX_HIP(hipMemcpyAsync(m, map.data(), map.size() * 4, hipMemcpyHostToDevice, st));
      FX_HIP(hipMemcpyAsync(c, coords.data(), coords.size() * 4, hipMemcpyHostToDevice, st));
      L.map = m;
      L.coords = c;
    }
    return FX_OK;
  }

  // seeded input faces; padding faces hold kPadding in both face arrays
  int seed(hipStream_t st) {
    const int s = L.side, per = (s + 1) * s * s;
    std::vector<float> in(size_t(L.faces) * 3), out(size_t(L.faces) * 3, 0.0f);
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < L.faces; ++i) {
        const int slot = i / per, j = i - slot * per;
        const int dx = s + (a == 0), dy = s + (a == 1);
        int c[3] = {j % dx, (j / dx) % dy, j / (dx * dy)};
        int b[3] = {0, 0, 0};
        if (L.sparse)
          for (int k = 0; k < 3; ++k) b[k] = coords[3 * slot + k];
        const bool padding = L.sparse && c[a] == s && b[a] != L.grid - 1;
        for (int k = 0; k < 3; ++k) c[k] += b[k] * s;
        in[size_t(a) * L.faces + i] = padding ? kPadding : seed_face(c[0], c[1], c[2], a, L.n, L.mode);
        out[size_t(a) * L.faces + i] = padding ? kPadding : 0.0f;
      }
    FX_HIP(hipMemcpyAsync(u_in, in.data(), in.size() * 4, hipMemcpyHostToDevice, st));
    FX_HIP(hipMemcpyAsync(u_out, out.data(), out.size() * 4, hipMemcpyHostToDevice, st));
    FX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    return FX_OK;
  }

  static dim3 blocks(int count) { return dim3((count + 255) / 256); }

  void smooth(int count, hipStream_t st) {
    for (int k = 0; k < count; ++k) {
      hipLaunchKernelGGL(k_jacobi, blocks(L.cells), dim3(256), 0, st, L, q[current], rhs, q[1 - current]);
      current = 1 - current;
    }
  }

  int cycle(std::vector<std::unique_ptr<Engine>>& levels, size_t level, hipStream_t st) {
    if (level == levels.size()) {
      smooth(128, st);
      return FX_OK;
    }
    smooth(4, st);
    hipLaunchKernelGGL(k_residual, blocks(L.cells), dim3(256), 0, st, L, q[current], rhs, after);
    Engine& c = *levels[level];
    FX_HIP(hipMemsetAsync(c.q[0], 0, uint64_t(c.L.cells) * 4, st));
    FX_HIP(hipMemsetAsync(c.q[1], 0, uint64_t(c.L.cells) * 4, st));
    c.current = 0;
    hipLaunchKernelGGL(k_restrict, blocks(c.L.cells), dim3(256), 0, st, L, c.L, after, c.rhs);
    if (int e = c.cycle(levels, level + 1, st)) return e;
    hipLaunchKernelGGL(k_prolong, blocks(L.cells), dim3(256), 0, st, L, c.L, c.q[c.current], q[current]);
    smooth(4, st);
    FX_LAUNCHED("brick projection cycle");
    return FX_OK;
  }

  // divergence, solve, gradient correction, divergence after; GPU ms
  int run(int iterations, int cycles, std::vector<std::unique_ptr<Engine>>& levels, hipStream_t st, double* ms) {
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    FX_HIP(hipEventRecord(e0, st));
    hipLaunchKernelGGL(k_divergence, blocks(L.cells), dim3(256), 0, st, L, u_in, rhs);
    if (cycles) {
      for (int c = 0; c < cycles; ++c)
        if (int e = cycle(levels, 0, st)) return e;
    } else {
      smooth(iterations, st);
    }
    hipLaunchKernelGGL(k_gradient, dim3((L.faces + 255) / 256, 3), dim3(256), 0, st, L, u_in, q[current], u_out);
    hipLaunchKernelGGL(k_divergence, blocks(L.cells), dim3(256), 0, st, L, u_out, after);
    FX_LAUNCHED("brick projection");
    FX_HIP(hipEventRecord(e1, st));
    FX_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    FX_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return FX_OK;
  }

  int download(const float* src, size_t count, std::vector<float>& out) {
    out.resize(count);
    FX_HIP(hipMemcpy(out.data(), src, count * 4, hipMemcpyDeviceToHost));
    return FX_OK;
  }
};

// bricks (x-fast order) with a solve cell within one cell of the brick: every
// solve face then has its canonical owner (projection.mm:379-385)
static void fine_topology(int n, int side, int mode, std::vector<int>& map, std::vector<int>& coords) {
  const int g = n / side;
  map.assign(size_t(g) * g * g, -1);
  coords.clear();
  for (int bz = 0; bz < g; ++bz)
    for (int by = 0; by < g; ++by)
      for (int bx = 0; bx < g; ++bx) {
        bool needed = false;
        for (int z = bz * side - 1; z <= (bz + 1) * side && !needed; ++z)
          for (int y = by * side - 1; y <= (by + 1) * side && !needed; ++y)
            for (int x = bx * side - 1; x <= (bx + 1) * side && !needed; ++x) needed = solve(x, y, z, n, mode);
        if (needed) {
          map[(size_t(bz) * g + by) * g + bx] = int(coords.size() / 3);
          coords.insert(coords.end(), {bx, by, bz});
        }
      }
}

// coarse-level bricks: those holding a solve cell of the level (projection.mm:396-400)
static void coarse_topology(int n, int side, int mode, std::vector<int>& map, std::vector<int>& coords) {
  const int g = n / side;
  map.assign(size_t(g) * g * g, -1);
  coords.clear();
  for (int bz = 0; bz < g; ++bz)
    for (int by = 0; by < g; ++by)
      for (int bx = 0; bx < g; ++bx) {
        bool active = false;
        for (int z = 0; z < side && !active; ++z)
          for (int y = 0; y < side && !active; ++y)
            for (int x = 0; x < side && !active; ++x) active = solve(bx * side + x, by * side + y, bz * side + z, n, mode);
        if (active) {
          map[(size_t(bz) * g + by) * g + bx] = int(coords.size() / 3);
          coords.insert(coords.end(), {bx, by, bz});
        }
      }
}

}  // namespace bp
}  // namespace fx

using namespace fx;
using namespace fx::bp;

extern "C" {

int fx_projection_compare(uint32_t n, uint32_t side, uint32_t iterations, uint32_t cycles, uint32_t mode,
                          uint64_t budget, float* faces, float* pressure, float* before, float* after,
                          fx_projection_report* r, void* stream) {
  FX_REQUIRE((n == 32 || n == 64 || n == 128) && (side == 8 || side == 16) && iterations <= 1024 && cycles <= 16 &&
                 mode <= 4 && budget > 0 && budget <= (uint64_t(1) << 30),
             FX_EINVAL, "Projection: n 32/64/128, brick 8/16, iterations 0..1024, mode 0..4, budget 1..1GiB");
  FX_REQUIRE(faces && pressure && before && after && r, FX_EINVAL, "null output");
  hipStream_t st = S(stream);
  *r = fx_projection_report{};
  r->resolution = n;
  r->side = side;
  r->iterations = iterations;
  r->mode = mode;
  r->cycles = cycles;
  r->levels = 1;
  const int N = int(n), s = int(side), g = N / s, M = int(mode);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> map, coords;
  fine_topology(N, s, M, map, coords);
  r->active_bricks = uint32_t(coords.size() / 3);
  r->topology_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  Engine sparse(N, s, M, true, map, coords, true);
  Engine dense(N, s, M, false, {}, {}, true);
  FX_REQUIRE(sparse.bytes <= budget && dense.bytes <= budget, FX_EINVAL, "Projection exceeds per-engine buffer budget");
  std::vector<std::unique_ptr<Engine>> sparse_levels, dense_levels;
  uint64_t sparse_total = sparse.bytes, dense_total = dense.bytes;
  if (cycles)
    for (int cn = N / 2; cn >= 8; cn /= 2) {
      const int cs = std::min(s, cn);
      std::vector<int> cm, cc;
      coarse_topology(cn, cs, M, cm, cc);
      auto se = std::make_unique<Engine>(cn, cs, M, true, cm, cc, false);
      auto de = std::make_unique<Engine>(cn, cs, M, false, std::vector<int>{}, std::vector<int>{}, false);
      FX_REQUIRE(sparse_total + se->bytes <= budget && dense_total + de->bytes <= budget, FX_EINVAL,
                 "Multigrid hierarchy exceeds per-engine budget");
      sparse_total += se->bytes;
      dense_total += de->bytes;
      sparse_levels.push_back(std::move(se));
      dense_levels.push_back(std::move(de));
    }
  (void)g;
  r->levels = uint32_t(sparse_levels.size()) + 1;
  r->sparse_bytes = sparse_total;
  r->dense_bytes = dense_total;
  r->sparse_hierarchy_bytes = sparse_total - sparse.bytes;
  r->dense_hierarchy_bytes = dense_total - dense.bytes;
  // every size and budget check passed: allocate, seed, run
  if (int e = sparse.allocate(st)) return e;
  if (int e = dense.allocate(st)) return e;
  for (auto& l : sparse_levels)
    if (int e = l->allocate(st)) return e;
  for (auto& l : dense_levels)
    if (int e = l->allocate(st)) return e;
  if (int e = sparse.seed(st)) return e;
  if (int e = dense.seed(st)) return e;
  if (int e = sparse.run(int(iterations), int(cycles), sparse_levels, st, &r->sparse_gpu_ms)) return e;
  if (int e = dense.run(int(iterations), int(cycles), dense_levels, st, &r->dense_gpu_ms)) return e;
  // read back
  std::vector<float> sq, srhs, safter, sfaces, dq, dafter, dfaces;
  if (int e = sparse.download(sparse.q[sparse.current], sparse.L.cells, sq)) return e;
  if (int e = sparse.download(sparse.rhs, sparse.L.cells, srhs)) return e;
  if (int e = sparse.download(sparse.after, sparse.L.cells, safter)) return e;
  if (int e = sparse.download(sparse.u_out, size_t(sparse.L.faces) * 3, sfaces)) return e;
  if (int e = dense.download(dense.q[dense.current], dense.L.cells, dq)) return e;
  if (int e = dense.download(dense.after, dense.L.cells, dafter)) return e;
  if (int e = dense.download(dense.u_out, size_t(dense.L.faces) * 3, dfaces)) return e;
  // dense-layout exports of the brick engine and the comparison figures
  auto sparse_cell = [&](int x, int y, int z) {
    const int slot = map[(size_t(z / s) * g + y / s) * g + x / s];
    return slot < 0 ? -1 : slot * s * s * s + ((z % s) * s + y % s) * s + x % s;
  };
  double before2 = 0, after2 = 0, res2 = 0;
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        const size_t i = (size_t(z) * N + y) * N + x;
        const int c = sparse_cell(x, y, z);
        pressure[i] = c < 0 ? 0.0f : sq[c];
        before[i] = c < 0 ? 0.0f : srhs[c];
        after[i] = c < 0 ? 0.0f : safter[c];
        FX_REQUIRE(std::isfinite(pressure[i]) && std::isfinite(after[i]), FX_ESTATE, "Nonfinite projection output");
        r->pressure_error = std::max(r->pressure_error, std::fabs(double(pressure[i]) - dq[i]));
        r->divergence_error = std::max(r->divergence_error, std::fabs(double(after[i]) - dafter[i]));
        if (solve(x, y, z, N, M)) {
          r->solve_cells++;
          before2 += double(before[i]) * before[i];
          after2 += double(after[i]) * after[i];
        }
      }
  // independent host residual rhs - L q from the exported canonical q
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        if (!solve(x, y, z, N, M)) continue;
        const size_t i = (size_t(z) * N + y) * N + x;
        const int c[3] = {x, y, z};
        double lap = 0;
        for (int a = 0; a < 3; ++a)
          for (int sg = -1; sg <= 1; sg += 2) {
            int v[3] = {c[0], c[1], c[2]};
            v[a] += sg;
            if (v[a] >= 0 && v[a] < N) lap += double(pressure[(size_t(v[2]) * N + v[1]) * N + v[0]]) - pressure[i];
          }
        const double res = before[i] - lap;
        res2 += res * res;
      }
  const double cells = std::max(1u, r->solve_cells);
  r->before_rms = std::sqrt(before2 / cells);
  r->after_rms = std::sqrt(after2 / cells);
  r->residual_rms = std::sqrt(res2 / cells);
  const size_t dc = size_t(N + 1) * N * N, sc = size_t(sparse.L.faces), per = size_t(s + 1) * s * s;
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; i < sc; ++i) {
      const size_t j = i % per;
      const int sx = s + (a == 0), sy = s + (a == 1);
      const int l[3] = {int(j % sx), int((j / sx) % sy), int(j / (size_t(sx) * sy))};
      const int* b = &coords[3 * (i / per)];
      if (l[a] == s && b[a] != g - 1)
        r->padding_error = std::max(r->padding_error, std::fabs(double(sfaces[a * sc + i]) - kPadding));
    }
    for (size_t i = 0; i < dc; ++i) {
      const int sx = N + (a == 0), sy = N + (a == 1);
      const int c[3] = {int(i % sx), int((i / sx) % sy), int(i / (size_t(sx) * sy))};
      int b[3] = {c[0] / s, c[1] / s, c[2] / s};
      b[a] = std::min(b[a], g - 1);
      const int slot = map[(size_t(b[2]) * g + b[1]) * g + b[0]];
      const int l[3] = {c[0] - b[0] * s, c[1] - b[1] * s, c[2] - b[2] * s};
      const float v =
          slot < 0 ? 0.0f : sfaces[a * sc + size_t(slot) * per + (size_t(l[2]) * (s + (a == 1)) + l[1]) * (s + (a == 0)) + l[0]];
      FX_REQUIRE(std::isfinite(v), FX_ESTATE, "Nonfinite projected face");
      faces[a * dc + i] = v;
      r->max_error = std::max(r->max_error, std::fabs(double(v) - dfaces[a * dc + i]));
      if (c[a] == 0 || c[a] == N) r->wall_error = std::max(r->wall_error, std::fabs(double(v)));
    }
  }
  return FX_OK;
}

}  // extern "C"
