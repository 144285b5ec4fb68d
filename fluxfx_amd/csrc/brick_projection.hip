// Unit-cell impulse projection and the coupled transport + projection +
// density step on 8^3 / 16^3 brick storage, each next to the same work on
// dense storage (the brick engine's diagnostic comparisons,
// native/projection.mm:14-462).
//
// Storage and canonical face ownership are those of brick_level.hpp; a
// brick engine stores the bricks with a solve cell within one cell of the
// brick (so every solve face has its owner), dense storage is the same code
// with one n^3 brick.  The solve region is analytic (a mode); the pressure is
// 0 outside it and on missing bricks, the outer walls are Neumann, and
// velocity faces outside the bricks read 0.  Jacobi and the V-cycle follow
// the reference engine: omega = 2/3, 4 + 4 sweeps per level, residual
// restricted as 4 x the 8-child mean (0.5 x the sum), cell-centred trilinear
// prolongation with clamped indices, 128 sweeps on the 8^3 level.
//
// A coupled step (projection.mm:347-366): faces self-advected (RK2, dt = 1
// voxel) and cut to the faces next to the solve region; divergence, V-cycles
// from q = 0, gradient correction; then density += emission on the solve
// region and density advected through the projected faces.  Density either
// shares the pressure bricks, or (hybrid) lives on bricks of its own next to
// dense pressure and velocity: a fixed support around the central half, or
// (adaptive) a support that only grows, extended before each density pass by
// a full-domain preflight that flags every brick a nonzero density can reach
// (projection.mm:126-137, 242-298).  Growth either reallocates the engine's
// block with its density (copying the block's prefix) or, pooled, keeps the
// density in buffers of their own whose capacity doubles when exceeded.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <vector>

#include "brick_level.hpp"

namespace fx {
namespace bp {

using namespace fx::brick;

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
        sum += cell_at(L, q, v[0], v[1], v[2]);
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
      if (inside(v[0], v[1], v[2], L.n)) lap += cell_at(L, q, v[0], v[1], v[2]) - q[i];
    }
  out[i] = rhs[i] - lap;
}

// an interior face with a solve cell on either side
__device__ __forceinline__ bool region_face(const Level& L, int a, const int c[3]) {
  int lo[3] = {c[0], c[1], c[2]};
  lo[a] -= 1;
  return c[a] > 0 && c[a] < L.n && (solve(c[0], c[1], c[2], L.n, L.mode) || solve(lo[0], lo[1], lo[2], L.n, L.mode));
}

// region faces: u - (q[c] - q[c - e_a]); other owned faces 0; padding faces
// untouched.  blockIdx.y = axis.
__global__ void k_gradient(Level L, const float* __restrict__ u, const float* __restrict__ q, float* __restrict__ out) {
  const int a = blockIdx.y;
  const int i = thread_id();
  if (i >= L.faces) return;
  int c[3];
  if (!face_coord(L, a, i, c)) return;
  float v = 0.0f;
  if (region_face(L, a, c)) {
    int lo[3] = {c[0], c[1], c[2]};
    lo[a] -= 1;
    v = u[a * L.faces + i] - (cell_at(L, q, c[0], c[1], c[2]) - cell_at(L, q, lo[0], lo[1], lo[2]));
  }
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
        for (int dx = 0; dx < 2; ++dx) sum += cell_at(F, fine, 2 * x + dx, 2 * y + dy, 2 * z + dz);
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
        const float w = corner_weight(dx, fx) * corner_weight(dy, fy) * corner_weight(dz, fz);
        sum = __builtin_fmaf(w, cell_at(C, coarse, qx, qy, qz), sum);
      }
  fine[i] += sum;
}

// faces self-advected (RK2 back-trace, dt = 1 voxel) on the region faces, 0
// on the other owned faces (projection.mm:101-109).  blockIdx.y = axis.
__global__ void k_transport_faces(Level L, const float* __restrict__ src, float* __restrict__ dst) {
  const int a = blockIdx.y;
  const int i = thread_id();
  if (i >= L.faces) return;
  int c[3];
  if (!face_coord(L, a, i, c)) return;
  float v = 0.0f;
  if (region_face(L, a, c)) {
    float p[3] = {c[0] + 0.5f, c[1] + 0.5f, c[2] + 0.5f}, d[3];
    p[a] = float(c[a]);
    departure(L, src, p, d);
    float q[3] = {d[0] - 0.5f, d[1] - 0.5f, d[2] - 0.5f};
    q[a] = d[a];
    v = sample_face(L, src, a, q[0], q[1], q[2]);
  }
  dst[a * L.faces + i] = v;
}

// density emitted into cell (x, y, z) at a step (projection.mm:110-116): a
// soft sphere of radius n/12 at (0.5, 0.5, 0.45) n, rate 0.1; schedule 1
// swings its centre in x, schedule 2 emits on even steps only
__device__ __forceinline__ float emission(int x, int y, int z, int n, int step, int schedule) {
  float cx = 0.5f * float(n);
  const float cy = 0.5f * float(n), cz = 0.45f * float(n);
  if (schedule == 1) cx += 0.05f * float(n) * sinf(0.5f * float(step));
  const float r = float(n) / 12.0f;
  const float dx = (x + 0.5f - cx) / r, dy = (y + 0.5f - cy) / r, dz = (z + 0.5f - cz) / r;
  const float rate = (schedule == 2 && (step & 1)) ? 0.0f : 0.1f;
  return rate * fmaxf(0.0f, 1.0f - (dx * dx + dy * dy + dz * dz));
}

__global__ void k_inject(Level D, const float* __restrict__ src, float* __restrict__ dst, int step, int schedule) {
  const int i = thread_id();
  if (i >= D.cells) return;
  int x, y, z;
  cell_coord(D, i, x, y, z);
  dst[i] = solve(x, y, z, D.n, D.mode) ? src[i] + emission(x, y, z, D.n, step, schedule) : 0.0f;
}

// cell-centred density traced back through the faces of V, sampled with the
// coordinate clamped to [0, n-1] (projection.mm:138-148)
__global__ void k_transport_density(Level D, Level V, const float* __restrict__ src, const float* __restrict__ u,
                                    float* __restrict__ dst) {
  const int i = thread_id();
  if (i >= D.cells) return;
  int x, y, z;
  cell_coord(D, i, x, y, z);
  float v = 0.0f;
  if (solve(x, y, z, D.n, D.mode)) {
    const float p[3] = {x + 0.5f, y + 0.5f, z + 0.5f};
    float d[3], q[3];
    departure(V, u, p, d);
    for (int k = 0; k < 3; ++k) q[k] = fminf(fmaxf(d[k] - 0.5f, 0.0f), float(D.n - 1));
    const int bx = int(floorf(q[0])), by = int(floorf(q[1])), bz = int(floorf(q[2]));
    const float fx = q[0] - bx, fy = q[1] - by, fz = q[2] - bz;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const float w = corner_weight(dx, fx) * corner_weight(dy, fy) * corner_weight(dz, fz);
          v = __builtin_fmaf(w, cell_at(D, src, min(bx + dx, D.n - 1), min(by + dy, D.n - 1), min(bz + dz, D.n - 1)), v);
        }
  }
  dst[i] = v;
}

// preflight over all n^3 cells: flag the brick of every cell whose density
// can be nonzero after the next inject + transport (an emitting cell, or a
// donor corner that holds density or emits) (projection.mm:126-137)
__global__ void k_required_density(Level D, Level V, const float* __restrict__ rho, const float* __restrict__ u,
                                   int* __restrict__ flags, int step, int schedule) {
  const int i = thread_id();
  const int n = D.n;
  if (i >= n * n * n) return;
  const int x = i % n, y = (i / n) % n, z = i / (n * n);
  bool needed = emission(x, y, z, n, step, schedule) > 0.0f;
  const float p[3] = {x + 0.5f, y + 0.5f, z + 0.5f};
  float d[3], q[3];
  departure(V, u, p, d);
  for (int k = 0; k < 3; ++k) q[k] = fminf(fmaxf(d[k] - 0.5f, 0.0f), float(n - 1));
  const int bx = int(floorf(q[0])), by = int(floorf(q[1])), bz = int(floorf(q[2]));
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int cx = min(bx + dx, n - 1), cy = min(by + dy, n - 1), cz = min(bz + dz, n - 1);
        needed = needed || cell_at(D, rho, cx, cy, cz) > 0.0f || emission(cx, cy, cz, n, step, schedule) > 0.0f;
      }
  if (needed) flags[((z / D.side) * D.grid + y / D.side) * D.grid + x / D.side] = 1;
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

// the density emitted at a step, in double on the host (the injected-mass figure)
static double emission_host(int x, int y, int z, int n, int step, int schedule) {
  if (schedule == 2 && step % 2) return 0.0;
  const double cx = 0.5 * n + (schedule == 1 ? 0.05 * n * std::sin(0.5 * step) : 0.0), r = double(n) / 12;
  const double dx = (x + 0.5 - cx) / r, dy = (y + 0.5 - 0.5 * n) / r, dz = (z + 0.5 - 0.45 * n) / r;
  return 0.1 * std::max(0.0, 1.0 - dx * dx - dy * dy - dz * dz);
}

static dim3 blocks(int64_t count) { return dim3(unsigned((count + 255) / 256)); }

static double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

struct Options {
  bool velocity = false;  // fine level: owns the face arrays
  bool coupled = false;   // owns density fields
  bool hybrid = false;    // density on bricks of its own next to dense pressure/velocity
  bool adaptive = false;  // hybrid density support grows by preflight
  bool pooled = false;    // adaptive density in buffers of their own with doubling capacity
  int seed_mode = 0;
};

struct Engine {
  Level L{}, D{};  // pressure/velocity level, density level
  Options o;
  int side = 8;    // brick side (density bricks of a hybrid engine)
  std::vector<int> map, coords, dmap, dcoords;
  // byte layout of the engine block: u_in, u_out, q0, q1, rhs, after, map,
  // coords [, density map, density coords, flags] [, rho0, rho1]
  uint64_t vel = 0, scal = 0, o_map = 0, o_coords = 0, o_dmap = 0, o_dcoords = 0, o_flags = 0, o_rho0 = 0;
  uint64_t dscal = 0;            // bytes of one density field in the block
  uint64_t bytes = 0, peak = 0;  // accounted bytes (block + pooled density), peak during growth
  uint64_t density_bytes = 0, storage = 0, preserved = 0, global_copy = 0;
  uint32_t capacity = 0, reallocs = 0, reuses = 0, expansions = 0;
  double preflight_ms = 0, density_topology_ms = 0;
  DeviceBuffer block, pool;
  float *u_in = nullptr, *u_out = nullptr, *q[2] = {nullptr, nullptr}, *rhs = nullptr, *after = nullptr;
  float* rho[2] = {nullptr, nullptr};
  int* flags = nullptr;
  int current = 0;

  Engine(int n, int side_, int mode, bool sparse, std::vector<int> bmap, std::vector<int> bcoords, Options opt)
      : o(opt), side(side_) {
    L.n = n;
    L.side = sparse ? side : n;
    L.grid = sparse ? n / side : 1;
    L.sparse = sparse ? 1 : 0;
    L.mode = mode;
    const int bricks = sparse ? int(bcoords.size() / 3) : 1;
    const int s = L.side;
    L.cells = bricks * s * s * s;
    L.faces = o.velocity ? bricks * (s + 1) * s * s : 0;
    map = std::move(bmap);
    coords = std::move(bcoords);
    vel = align256(uint64_t(L.faces) * 3 * 4);
    scal = align256(uint64_t(L.cells) * 4);
    o_map = 2 * vel + 4 * scal;
    o_coords = o_map + align256(map.size() * 4);
    bytes = o_coords + align256(coords.size() * 4);
    const uint64_t fixed = bytes;
    D = L;
    if (o.hybrid) {
      const auto t0 = std::chrono::steady_clock::now();
      const int g = n / side;
      // fixed support: the central half plus one brick each side; adaptive:
      // the bricks holding the centre cell (projection.mm:183-192)
      const int lo = o.adaptive ? n / 2 : n / 4 - side, hi = o.adaptive ? n / 2 + 1 : 3 * n / 4 + side;
      dmap.assign(size_t(g) * g * g, -1);
      auto out = [&](int b) { return (b + 1) * side <= lo || b * side >= hi; };
      for (int z = 0; z < g; ++z)
        for (int y = 0; y < g; ++y)
          for (int x = 0; x < g; ++x) {
            if (out(x) || out(y) || out(z)) continue;
            dmap[(size_t(z) * g + y) * g + x] = int(dcoords.size() / 3);
            dcoords.insert(dcoords.end(), {x, y, z});
          }
      D.side = side;
      D.grid = g;
      D.sparse = 1;
      D.mode = 0;
      D.cells = int(dcoords.size() / 3) * side * side * side;
      D.faces = 0;
      o_dmap = bytes;
      o_dcoords = bytes + align256(dmap.size() * 4);
      // the adaptive coordinate table is sized for every brick up front
      bytes = o_dcoords + align256((o.adaptive ? dmap.size() : dcoords.size() / 3) * 12);
      if (o.adaptive) {
        o_flags = bytes;
        bytes += align256(dmap.size() * 4);
      }
      density_topology_ms = ms_since(t0);
    }
    dscal = align256(uint64_t(D.cells) * 4);
    o_rho0 = bytes;
    if (o.coupled) bytes += 2 * dscal;
    density_bytes = o.coupled ? bytes - fixed : 0;
    if (o.pooled) {
      capacity = uint32_t(dcoords.size() / 3);
      storage = 2 * dscal;
    }
    peak = bytes;
  }

  // device pointers from the block (and pool) base addresses
  void bind() {
    char* b = static_cast<char*>(block.p);
    u_in = reinterpret_cast<float*>(b);
    u_out = reinterpret_cast<float*>(b + vel);
    q[0] = reinterpret_cast<float*>(b + 2 * vel);
    q[1] = reinterpret_cast<float*>(b + 2 * vel + scal);
    rhs = reinterpret_cast<float*>(b + 2 * vel + 2 * scal);
    after = reinterpret_cast<float*>(b + 2 * vel + 3 * scal);
    L.map = L.sparse ? reinterpret_cast<int*>(b + o_map) : nullptr;
    L.coords = L.sparse ? reinterpret_cast<int*>(b + o_coords) : nullptr;
    if (o.hybrid) {
      D.map = reinterpret_cast<int*>(b + o_dmap);
      D.coords = reinterpret_cast<int*>(b + o_dcoords);
      flags = o.adaptive ? reinterpret_cast<int*>(b + o_flags) : nullptr;
    } else {
      D.map = L.map;
      D.coords = L.coords;
    }
    if (!o.coupled) return;
    if (o.pooled) {
      rho[0] = static_cast<float*>(pool.p);
      rho[1] = reinterpret_cast<float*>(static_cast<char*>(pool.p) + storage / 2);
    } else {
      rho[0] = reinterpret_cast<float*>(b + o_rho0);
      rho[1] = reinterpret_cast<float*>(b + o_rho0 + dscal);
    }
  }

  int upload_tables(hipStream_t st) {
    char* b = static_cast<char*>(block.p);
    if (L.sparse) {
      FX_HIP(hipMemcpyAsync(b + o_map, map.data(), map.size() * 4, hipMemcpyHostToDevice, st));
      FX_HIP(hipMemcpyAsync(b + o_coords, coords.data(), coords.size() * 4, hipMemcpyHostToDevice, st));
    }
    if (o.hybrid) {
      FX_HIP(hipMemcpyAsync(b + o_dmap, dmap.data(), dmap.size() * 4, hipMemcpyHostToDevice, st));
      FX_HIP(hipMemcpyAsync(b + o_dcoords, dcoords.data(), dcoords.size() * 4, hipMemcpyHostToDevice, st));
    }
    return FX_OK;
  }

  int allocate(hipStream_t st) {
    FX_HIP(hipMalloc(&block.p, bytes - storage));
    FX_HIP(hipMemsetAsync(block.p, 0, bytes - storage, st));
    if (o.pooled) {
      FX_HIP(hipMalloc(&pool.p, storage));
      FX_HIP(hipMemsetAsync(pool.p, 0, storage, st));
    }
    bind();
    if (int e = upload_tables(st)) return e;
    FX_HIP(hipStreamSynchronize(st));  // the host tables may change before their next upload
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
        in[size_t(a) * L.faces + i] = padding ? kPadding : seed_face(c[0], c[1], c[2], a, L.n, o.seed_mode);
        out[size_t(a) * L.faces + i] = padding ? kPadding : 0.0f;
      }
    FX_HIP(hipMemcpyAsync(u_in, in.data(), in.size() * 4, hipMemcpyHostToDevice, st));
    FX_HIP(hipMemcpyAsync(u_out, out.data(), out.size() * 4, hipMemcpyHostToDevice, st));
    FX_HIP(hipStreamSynchronize(st));  // the host vectors go out of scope
    return FX_OK;
  }

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
    FX_HIP(hipMemsetAsync(c.q[0], 0, 2 * c.scal, st));  // q0 and q1 are adjacent
    c.current = 0;
    hipLaunchKernelGGL(k_restrict, blocks(c.L.cells), dim3(256), 0, st, L, c.L, after, c.rhs);
    if (int e = c.cycle(levels, level + 1, st)) return e;
    hipLaunchKernelGGL(k_prolong, blocks(L.cells), dim3(256), 0, st, L, c.L, c.q[c.current], q[current]);
    smooth(4, st);
    FX_LAUNCHED("brick projection cycle");
    return FX_OK;
  }

  // divergence, solve, gradient correction, divergence after
  int project(int iterations, int cycles, std::vector<std::unique_ptr<Engine>>& levels, hipStream_t st) {
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
    return FX_OK;
  }

  // extend the density support to every brick the next density pass can
  // reach; new bricks append slots, so the stored prefix keeps its layout
  int expand(int step, int schedule, uint64_t budget, hipStream_t st) {
    const auto t0 = std::chrono::steady_clock::now();
    FX_HIP(hipMemsetAsync(flags, 0, dmap.size() * 4, st));
    hipLaunchKernelGGL(k_required_density, blocks(int64_t(L.n) * L.n * L.n), dim3(256), 0, st, D, L, rho[0], u_out,
                       flags, step, schedule);
    FX_LAUNCHED("density preflight");
    std::vector<int> flagged(dmap.size());
    FX_HIP(hipMemcpyAsync(flagged.data(), flags, flagged.size() * 4, hipMemcpyDeviceToHost, st));
    FX_HIP(hipStreamSynchronize(st));
    std::vector<int> next_map = dmap, next_coords = dcoords;
    const int g = D.grid;
    for (int z = 0; z < g; ++z)
      for (int y = 0; y < g; ++y)
        for (int x = 0; x < g; ++x) {
          const size_t i = (size_t(z) * g + y) * g + x;
          if (flagged[i] && next_map[i] < 0) {
            next_map[i] = int(next_coords.size() / 3);
            next_coords.insert(next_coords.end(), {x, y, z});
          }
        }
    const uint32_t count = uint32_t(next_coords.size() / 3), cube = uint32_t(side) * side * side;
    if (count != dcoords.size() / 3) {
      const uint64_t used = uint64_t(D.cells) * 4;
      std::vector<char> kept(used), copied(used);
      if (o.pooled) {
        if (count > capacity) {
          uint32_t cap = capacity;
          while (cap < count) cap *= 2;
          cap = std::min<uint32_t>(cap, uint32_t(dmap.size()));
          const uint64_t field = uint64_t(cap) * cube * 4, fresh = 2 * field;
          FX_REQUIRE(bytes + fresh <= budget, FX_EINVAL, "Density capacity growth exceeds transient buffer budget");
          DeviceBuffer grown;
          FX_HIP(hipMalloc(&grown.p, fresh));
          FX_HIP(hipMemsetAsync(grown.p, 0, fresh, st));
          FX_HIP(hipMemcpyAsync(grown.p, rho[0], used, hipMemcpyDeviceToDevice, st));
          FX_HIP(hipMemcpyAsync(kept.data(), rho[0], used, hipMemcpyDeviceToHost, st));
          FX_HIP(hipMemcpyAsync(copied.data(), grown.p, used, hipMemcpyDeviceToHost, st));
          FX_HIP(hipStreamSynchronize(st));
          FX_REQUIRE(std::memcmp(kept.data(), copied.data(), used) == 0, FX_ESTATE,
                     "Density capacity preservation failed");
          preserved += used;
          peak = std::max(peak, bytes + fresh);
          bytes += fresh - storage;
          density_bytes += fresh - storage;
          storage = fresh;
          std::swap(pool.p, grown.p);
          capacity = cap;
          ++reallocs;
        } else {
          ++reuses;
        }
      } else {
        const uint64_t cells = uint64_t(count) * cube, field = align256(cells * 4), fresh = o_rho0 + 2 * field;
        // the old and the new block coexist while the prefix is copied
        FX_REQUIRE(bytes + fresh <= budget, FX_EINVAL, "Adaptive density growth exceeds transient buffer budget");
        DeviceBuffer grown;
        FX_HIP(hipMalloc(&grown.p, fresh));
        FX_HIP(hipMemsetAsync(grown.p, 0, fresh, st));
        FX_HIP(hipMemcpyAsync(grown.p, block.p, o_rho0 + used, hipMemcpyDeviceToDevice, st));
        FX_HIP(hipMemcpyAsync(kept.data(), rho[0], used, hipMemcpyDeviceToHost, st));
        FX_HIP(hipMemcpyAsync(copied.data(), static_cast<char*>(grown.p) + o_rho0, used, hipMemcpyDeviceToHost, st));
        FX_HIP(hipStreamSynchronize(st));
        FX_REQUIRE(std::memcmp(kept.data(), copied.data(), used) == 0, FX_ESTATE, "Density preservation check failed");
        preserved += used;
        global_copy += o_rho0;
        peak = std::max(peak, bytes + fresh);
        density_bytes += fresh - bytes;
        bytes = fresh;
        dscal = field;
        std::swap(block.p, grown.p);
      }
      D.cells = int(count * cube);
      dmap = std::move(next_map);
      dcoords = std::move(next_coords);
      bind();
      if (int e = upload_tables(st)) return e;
      FX_HIP(hipStreamSynchronize(st));
      ++expansions;
    }
    preflight_ms += ms_since(t0);
    return FX_OK;
  }

  // one coupled step (projection.mm:347-366)
  int coupled_step(int step, int schedule, int cycles, std::vector<std::unique_ptr<Engine>>& levels, uint64_t budget,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_transport_faces, dim3((L.faces + 255) / 256, 3), dim3(256), 0, st, L, u_in, u_out);
    FX_HIP(hipMemcpyAsync(u_in, u_out, vel, hipMemcpyDeviceToDevice, st));
    FX_HIP(hipMemsetAsync(q[0], 0, 2 * scal, st));
    current = 0;
    if (int e = project(0, cycles, levels, st)) return e;
    if (o.adaptive)
      if (int e = expand(step, schedule, budget, st)) return e;
    hipLaunchKernelGGL(k_inject, blocks(D.cells), dim3(256), 0, st, D, rho[0], rho[1], step, schedule);
    hipLaunchKernelGGL(k_transport_density, blocks(D.cells), dim3(256), 0, st, D, L, rho[1], u_out, rho[0]);
    FX_HIP(hipMemcpyAsync(u_in, u_out, vel, hipMemcpyDeviceToDevice, st));
    FX_LAUNCHED("coupled step");
    return FX_OK;
  }

  int download(const float* src, size_t count, std::vector<float>& out) {
    out.resize(count);
    FX_HIP(hipMemcpy(out.data(), src, count * 4, hipMemcpyDeviceToHost));
    return FX_OK;
  }

  // stored index of density cell (x, y, z), or -1
  int density_index(int x, int y, int z) const {
    if (!D.sparse) return (z * D.n + y) * D.n + x;
    const std::vector<int>& m = o.hybrid ? dmap : map;
    const int slot = m[(size_t(z / side) * D.grid + y / side) * D.grid + x / side];
    return slot < 0 ? -1 : slot * side * side * side + ((z % side) * side + y % side) * side + x % side;
  }
};

// bricks with a solve cell within one cell of the brick: every solve face
// then has its canonical owner (projection.mm:379-385)
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

struct Request {
  uint32_t n, side, iterations, cycles, mode, steps, schedule;
  bool dense_full, global, hybrid, adaptive, pooled;
  uint64_t budget;
};

// both engines, the comparison and the dense-layout exports of the brick engine
static int compare(const Request& q, float* faces, float* pressure, float* before, float* after, float* density,
                   fx_coupled_report* out, hipStream_t st) {
  FX_REQUIRE(!q.pooled || q.adaptive, FX_EINVAL, "Density capacity requires adaptive mode");
  FX_REQUIRE(!q.adaptive || q.hybrid, FX_EINVAL, "Adaptive requires hybrid");
  FX_REQUIRE(!q.hybrid || (q.global && q.dense_full && q.steps), FX_EINVAL, "Hybrid requires global coupled comparison");
  FX_REQUIRE((q.n == 32 || q.n == 64 || q.n == 128) && (q.side == 8 || q.side == 16) && q.iterations <= 1024 &&
                 q.cycles <= 16 && q.steps <= 16 && q.schedule <= 2 && !(q.steps && !q.cycles) && q.mode <= 4 &&
                 q.budget > 0 && q.budget <= (uint64_t(1) << 30),
             FX_EINVAL, "Projection: n 32/64/128, brick 8/16, iterations 0..1024, mode 0..4, budget 1..1GiB");
  FX_REQUIRE(faces && pressure && before && after && out && (density || !q.steps), FX_EINVAL, "null output");
  *out = fx_coupled_report{};
  fx_projection_report* r = &out->projection;
  r->resolution = q.n;
  r->side = q.side;
  r->iterations = q.iterations;
  r->mode = q.mode;
  r->cycles = q.cycles;
  r->levels = 1;
  out->steps = q.steps;
  out->schedule = q.schedule;
  out->dense_full = q.dense_full;
  out->global_support = q.global;
  out->hybrid = q.hybrid;
  out->adaptive = q.adaptive;
  out->pooled = q.pooled;
  const int N = int(q.n), s = int(q.side), g = N / s, M = int(q.mode);
  const int solve_mode = q.global ? 0 : M, dense_mode = q.dense_full ? 0 : solve_mode;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> map, coords;
  fine_topology(N, s, solve_mode, map, coords);
  r->active_bricks = uint32_t(coords.size() / 3);
  r->topology_ms = ms_since(t0);
  Options so, dopt;
  so.velocity = dopt.velocity = true;
  so.coupled = dopt.coupled = q.steps > 0;
  so.seed_mode = dopt.seed_mode = M;
  so.hybrid = q.hybrid;
  so.adaptive = q.adaptive;
  so.pooled = q.pooled;
  Engine sparse(N, s, solve_mode, !q.hybrid, map, coords, so);
  Engine dense(N, s, dense_mode, false, {}, {}, dopt);
  FX_REQUIRE(sparse.bytes <= q.budget && dense.bytes <= q.budget, FX_EINVAL,
             "Projection exceeds per-engine buffer budget");
  std::vector<std::unique_ptr<Engine>> sparse_levels, dense_levels;
  uint64_t sparse_total = sparse.bytes, dense_total = dense.bytes;
  if (q.cycles)
    for (int cn = N / 2; cn >= 8; cn /= 2) {
      const int cs = std::min(s, cn);
      std::vector<int> cm, cc;
      coarse_topology(cn, cs, solve_mode, cm, cc);
      auto se = std::make_unique<Engine>(cn, cs, solve_mode, !q.hybrid, cm, cc, Options{});
      auto de = std::make_unique<Engine>(cn, cs, dense_mode, false, std::vector<int>{}, std::vector<int>{}, Options{});
      FX_REQUIRE(sparse_total + se->bytes <= q.budget && dense_total + de->bytes <= q.budget, FX_EINVAL,
                 "Multigrid hierarchy exceeds per-engine budget");
      sparse_total += se->bytes;
      dense_total += de->bytes;
      sparse_levels.push_back(std::move(se));
      dense_levels.push_back(std::move(de));
    }
  r->levels = uint32_t(sparse_levels.size()) + 1;
  r->sparse_hierarchy_bytes = sparse_total - sparse.bytes;
  r->dense_hierarchy_bytes = dense_total - dense.bytes;
  r->sparse_bytes = sparse_total;
  r->dense_bytes = dense_total;
  out->density_topology_ms = sparse.density_topology_ms;
  // every size and budget check passed: allocate, seed, run
  if (int e = sparse.allocate(st)) return e;
  if (int e = dense.allocate(st)) return e;
  for (auto& l : sparse_levels)
    if (int e = l->allocate(st)) return e;
  for (auto& l : dense_levels)
    if (int e = l->allocate(st)) return e;
  if (int e = sparse.seed(st)) return e;
  if (int e = dense.seed(st)) return e;
  auto run = [&](Engine& e, std::vector<std::unique_ptr<Engine>>& levels, uint64_t budget, double* gpu_ms,
                 double* wall_ms) -> int {
    const auto w0 = std::chrono::steady_clock::now();
    hipEvent_t e0, e1;
    FX_HIP(hipEventCreate(&e0));
    FX_HIP(hipEventCreate(&e1));
    int err = FX_OK;
    if (hipEventRecord(e0, st) != hipSuccess) err = fail_hip(hipGetLastError(), "hipEventRecord");
    if (!err && q.steps) {
      for (uint32_t k = 0; k < q.steps && !err; ++k)
        err = e.coupled_step(int(k), int(q.schedule), int(q.cycles), levels, budget, st);
    } else if (!err) {
      err = e.project(int(q.iterations), int(q.cycles), levels, st);
    }
    float t = 0.0f;
    if (!err) {
      hipError_t h = hipEventRecord(e1, st);
      if (h == hipSuccess) h = hipEventSynchronize(e1);
      if (h == hipSuccess) h = hipEventElapsedTime(&t, e0, e1);
      if (h != hipSuccess) err = fail_hip(h, "brick comparison timing");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gpu_ms = t;
    *wall_ms = ms_since(w0);
    return err;
  };
  if (int e = run(sparse, sparse_levels, q.budget - r->sparse_hierarchy_bytes, &r->sparse_gpu_ms, &out->sparse_wall_ms))
    return e;
  if (int e = run(dense, dense_levels, q.budget - r->dense_hierarchy_bytes, &r->dense_gpu_ms, &out->dense_wall_ms))
    return e;
  r->sparse_bytes = sparse.bytes + r->sparse_hierarchy_bytes;
  out->peak_sparse_bytes = sparse.peak + r->sparse_hierarchy_bytes;
  out->density_bytes = sparse.density_bytes;
  out->density_cells = q.steps ? uint64_t(sparse.D.cells) : 0;
  out->density_bricks = q.steps ? uint32_t(sparse.D.cells / (s * s * s)) : 0;
  out->density_capacity = sparse.capacity;
  out->density_reallocations = sparse.reallocs;
  out->density_reuses = sparse.reuses;
  out->preserved_bytes = sparse.preserved;
  out->global_copy_bytes = sparse.global_copy;
  out->expansions = sparse.expansions;
  out->preflight_ms = sparse.preflight_ms;
  // read back
  std::vector<float> sq, srhs, safter, sfaces, srho, dq, dafter, dfaces, drho;
  if (int e = sparse.download(sparse.q[sparse.current], sparse.L.cells, sq)) return e;
  if (int e = sparse.download(sparse.rhs, sparse.L.cells, srhs)) return e;
  if (int e = sparse.download(sparse.after, sparse.L.cells, safter)) return e;
  if (int e = sparse.download(sparse.u_out, size_t(sparse.L.faces) * 3, sfaces)) return e;
  if (int e = dense.download(dense.q[dense.current], dense.L.cells, dq)) return e;
  if (int e = dense.download(dense.after, dense.L.cells, dafter)) return e;
  if (int e = dense.download(dense.u_out, size_t(dense.L.faces) * 3, dfaces)) return e;
  if (q.steps) {
    if (int e = sparse.download(sparse.rho[0], sparse.D.cells, srho)) return e;
    if (int e = dense.download(dense.rho[0], dense.D.cells, drho)) return e;
  }
  // dense-layout exports of the brick engine and the comparison figures
  auto sparse_cell = [&](int x, int y, int z) {
    if (q.hybrid) return (z * N + y) * N + x;
    const int slot = map[(size_t(z / s) * g + y / s) * g + x / s];
    return slot < 0 ? -1 : slot * s * s * s + ((z % s) * s + y % s) * s + x % s;
  };
  double before2 = 0, after2 = 0, res2 = 0, dense_after2 = 0;
  uint32_t dense_cells = 0;
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
        if (solve(x, y, z, N, dense_mode)) {
          dense_after2 += double(dafter[i]) * dafter[i];
          ++dense_cells;
        }
        if (q.steps) {
          const int di = sparse.density_index(x, y, z);
          const float v = di < 0 ? 0.0f : srho[di], w = drho[i];
          FX_REQUIRE(std::isfinite(v) && std::isfinite(w) && v >= 0.0f && w >= 0.0f, FX_ESTATE,
                     "Invalid coupled density");
          density[i] = v;
          out->density_error = std::max(out->density_error, std::fabs(double(v) - w));
          out->sparse_mass += v;
          out->dense_mass += w;
          if (solve(x, y, z, N, solve_mode))
            for (uint32_t k = 0; k < q.steps; ++k) out->injected_mass += emission_host(x, y, z, N, int(k), int(q.schedule));
        }
        if (solve(x, y, z, N, solve_mode)) {
          r->solve_cells++;
          before2 += double(before[i]) * before[i];
          after2 += double(after[i]) * after[i];
        }
      }
  // independent host residual rhs - L q from the exported canonical q
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        if (!solve(x, y, z, N, solve_mode)) continue;
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
  out->dense_after_rms = std::sqrt(dense_after2 / std::max(1u, dense_cells));
  const size_t dc = size_t(N + 1) * N * N, sc = size_t(sparse.L.faces), per = size_t(s + 1) * s * s;
  for (int a = 0; a < 3; ++a) {
    for (size_t i = 0; !q.hybrid && i < sc; ++i) {
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
      float v;
      if (q.hybrid) {
        v = sfaces[a * dc + i];
      } else {
        int b[3] = {c[0] / s, c[1] / s, c[2] / s};
        b[a] = std::min(b[a], g - 1);
        const int slot = map[(size_t(b[2]) * g + b[1]) * g + b[0]];
        const int l[3] = {c[0] - b[0] * s, c[1] - b[1] * s, c[2] - b[2] * s};
        v = slot < 0 ? 0.0f
                     : sfaces[a * sc + size_t(slot) * per + (size_t(l[2]) * (s + (a == 1)) + l[1]) * (s + (a == 0)) + l[0]];
      }
      FX_REQUIRE(std::isfinite(v), FX_ESTATE, "Nonfinite projected face");
      faces[a * dc + i] = v;
      r->max_error = std::max(r->max_error, std::fabs(double(v) - dfaces[a * dc + i]));
      if (c[a] == 0 || c[a] == N) r->wall_error = std::max(r->wall_error, std::fabs(double(v)));
    }
  }
  return FX_OK;
}

}  // namespace bp
}  // namespace fx

using namespace fx;
using namespace fx::bp;

extern "C" {

int fx_projection_compare(uint32_t n, uint32_t side, uint32_t iterations, uint32_t cycles, uint32_t mode,
                          uint64_t budget, float* faces, float* pressure, float* before, float* after,
                          fx_projection_report* r, void* stream) {
  FX_REQUIRE(r, FX_EINVAL, "null output");
  fx_coupled_report full{};
  const Request q{n, side, iterations, cycles, mode, 0, 0, false, false, false, false, false, budget};
  const int e = compare(q, faces, pressure, before, after, nullptr, &full, S(stream));
  *r = full.projection;
  return e;
}

int fx_coupled_compare(uint32_t n, uint32_t side, uint32_t steps, uint32_t cycles, uint32_t mode, uint32_t schedule,
                       uint32_t options, uint64_t budget, float* faces, float* pressure, float* before, float* after,
                       float* density, fx_coupled_report* r, void* stream) {
  FX_REQUIRE(r, FX_EINVAL, "null output");
  FX_REQUIRE(steps >= 1 && steps <= 16, FX_EINVAL, "Use 1..16 coupled steps");
  FX_REQUIRE(cycles >= 1 && cycles <= 16, FX_EINVAL, "Use 1..16 multigrid cycles");
  FX_REQUIRE((options & ~uint32_t(31)) == 0, FX_EINVAL, "unknown coupled option bits");
  const Request q{n,
                  side,
                  0,
                  cycles,
                  mode,
                  steps,
                  schedule,
                  (options & FX_COUPLED_DENSE_FULL) != 0,
                  (options & FX_COUPLED_GLOBAL) != 0,
                  (options & FX_COUPLED_HYBRID) != 0,
                  (options & FX_COUPLED_ADAPTIVE) != 0,
                  (options & FX_COUPLED_POOLED) != 0,
                  budget};
  return compare(q, faces, pressure, before, after, density, r, S(stream));
}

}  // extern "C"
