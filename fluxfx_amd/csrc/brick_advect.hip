// Advection diagnostics on brick storage, each next to the same work on
// dense storage (native/transport.mm, native/mac.mm):
//
// * constant-displacement scalar transport: a centred cube of ones moved by
//   (dx, dy, dz) voxels per step, each step a trilinear sample at
//   cell - displacement with the integer indices clamped to the box; the
//   displacement is the same for every cell, so the corner offset and the
//   weights are computed once on the host and every cell does 8 loads and 7
//   lerps;
// * MAC self-advection with dt = 1 voxel: each stored (canonically owned)
//   face traces back through the staggered velocity by RK2 and samples its
//   own component, coordinates clamped to the face array.
//
// Missing bricks read 0; the brick set covers every cell the field can reach
// in the requested steps, so brick and dense results agree exactly.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "brick_level.hpp"

namespace fx {
namespace ba {

using namespace fx::brick;

__device__ __forceinline__ float lerp(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }

// cell value with the indices clamped to the box; missing bricks read 0
__device__ __forceinline__ float clamped(const Level& L, const float* f, int x, int y, int z) {
  const int i = cell_index(L, min(max(x, 0), L.n - 1), min(max(y, 0), L.n - 1), min(max(z, 0), L.n - 1));
  return i < 0 ? 0.0f : f[i];
}

// dst = trilinear sample of src at cell + base + frac (transport.mm:27-36)
__global__ void k_translate(Level L, const float* __restrict__ src, float* __restrict__ dst, int bx, int by, int bz,
                            float fx, float fy, float fz) {
  const int i = thread_id();
  if (i >= L.cells) return;
  int x, y, z;
  cell_coord(L, i, x, y, z);
  x += bx;
  y += by;
  z += bz;
  const float a = lerp(clamped(L, src, x, y, z), clamped(L, src, x + 1, y, z), fx);
  const float b = lerp(clamped(L, src, x, y + 1, z), clamped(L, src, x + 1, y + 1, z), fx);
  const float c = lerp(clamped(L, src, x, y, z + 1), clamped(L, src, x + 1, y, z + 1), fx);
  const float d = lerp(clamped(L, src, x, y + 1, z + 1), clamped(L, src, x + 1, y + 1, z + 1), fx);
  dst[i] = lerp(lerp(a, b, fy), lerp(c, d, fy), fz);
}

// one self-advection step of every owned face (mac.mm:40-45); blockIdx.y = axis
__global__ void k_mac_advect(Level L, const float* __restrict__ src, float* __restrict__ dst) {
  const int a = blockIdx.y;
  const int i = thread_id();
  if (i >= L.faces) return;
  int c[3];
  if (!face_coord(L, a, i, c)) return;  // padding: owned by the next brick
  float p[3] = {c[0] + 0.5f, c[1] + 0.5f, c[2] + 0.5f}, d[3];
  p[a] = float(c[a]);
  departure(L, src, p, d);
  float q[3] = {d[0] - 0.5f, d[1] - 0.5f, d[2] - 0.5f};
  q[a] = d[a];
  dst[a * L.faces + i] = sample_face(L, src, a, q[0], q[1], q[2]);
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
constexpr float kPadding = 12345.0f;

static double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

// a brick (or dense) field pair with its tables in one allocation
struct Pair {
  Level L{};
  std::vector<int> map, coords;
  uint64_t stride = 0, o_map = 0, o_coords = 0, bytes = 0;
  DeviceBuffer mem;
  float* f[2] = {nullptr, nullptr};
  int cur = 0;

  // faces: each field holds the three face arrays, else one cell array
  Pair(int n, int side, bool sparse, std::vector<int> bmap, std::vector<int> bcoords, bool faces) {
    L.n = n;
    L.side = sparse ? side : n;
    L.grid = sparse ? n / side : 1;
    L.sparse = sparse ? 1 : 0;
    L.mode = 0;
    const int bricks = sparse ? int(bcoords.size() / 3) : 1, s = L.side;
    L.cells = bricks * s * s * s;
    L.faces = faces ? bricks * (s + 1) * s * s : 0;
    map = std::move(bmap);
    coords = std::move(bcoords);
    stride = align256(faces ? uint64_t(L.faces) * 3 * 4 : uint64_t(L.cells) * 4);
    o_map = 2 * stride;
    o_coords = o_map + align256(map.size() * 4);
    bytes = o_coords + align256(coords.size() * 4);
  }

  int allocate(const std::vector<float>& init, const std::vector<float>& other, hipStream_t st) {
    FX_HIP(hipMalloc(&mem.p, bytes));
    char* b = static_cast<char*>(mem.p);
    f[0] = reinterpret_cast<float*>(b);
    f[1] = reinterpret_cast<float*>(b + stride);
    FX_HIP(hipMemcpyAsync(f[0], init.data(), init.size() * 4, hipMemcpyHostToDevice, st));
    FX_HIP(hipMemcpyAsync(f[1], other.data(), other.size() * 4, hipMemcpyHostToDevice, st));
    if (L.sparse) {
      FX_HIP(hipMemcpyAsync(b + o_map, map.data(), map.size() * 4, hipMemcpyHostToDevice, st));
      FX_HIP(hipMemcpyAsync(b + o_coords, coords.data(), coords.size() * 4, hipMemcpyHostToDevice, st));
      L.map = reinterpret_cast<int*>(b + o_map);
      L.coords = reinterpret_cast<int*>(b + o_coords);
    }
    FX_HIP(hipStreamSynchronize(st));  // the host vectors may go out of scope
    return FX_OK;
  }

  // global cell of stored cell i
  void cell(size_t i, int c[3]) const {
    const int s = L.side, per = s * s * s, j = int(i % per);
    c[0] = j % s;
    c[1] = (j / s) % s;
    c[2] = j / (s * s);
    if (L.sparse)
      for (int k = 0; k < 3; ++k) c[k] += coords[3 * (i / per) + k] * s;
  }

  // global face of stored face i of axis a; false for padding
  bool face(int a, size_t i, int c[3]) const {
    const int s = L.side, per = (s + 1) * s * s, j = int(i % per);
    const int dx = s + (a == 0), dy = s + (a == 1);
    c[0] = j % dx;
    c[1] = (j / dx) % dy;
    c[2] = j / (dx * dy);
    if (!L.sparse) return true;
    const int* b = &coords[3 * (i / per)];
    if (c[a] == s && b[a] != L.grid - 1) return false;
    for (int k = 0; k < 3; ++k) c[k] += b[k] * s;
    return true;
  }

  // stored index of cell (x, y, z) / face (x, y, z) of axis a, or -1
  long cell_index(int x, int y, int z) const {
    if (!L.sparse) return (long(z) * L.n + y) * L.n + x;
    const int s = L.side, slot = map[(size_t(z / s) * L.grid + y / s) * L.grid + x / s];
    return slot < 0 ? -1 : long(slot) * s * s * s + ((z % s) * s + y % s) * s + x % s;
  }
  long face_index(int a, int x, int y, int z) const {
    if (!L.sparse) return long(a) * L.faces + long(dense_face(L.n, a, x, y, z));
    const int s = L.side;
    int b[3] = {x / s, y / s, z / s};
    b[a] = std::min(b[a], L.grid - 1);
    const int slot = map[(size_t(b[2]) * L.grid + b[1]) * L.grid + b[0]];
    if (slot < 0) return -1;
    const int l[3] = {x - b[0] * s, y - b[1] * s, z - b[2] * s};
    return long(a) * L.faces + long(slot) * (s + 1) * s * s + (long(l[2]) * (s + (a == 1)) + l[1]) * (s + (a == 0)) + l[0];
  }

  int download(std::vector<float>& out) const {
    out.resize(L.faces ? size_t(L.faces) * 3 : size_t(L.cells));
    FX_HIP(hipMemcpy(out.data(), f[cur], out.size() * 4, hipMemcpyDeviceToHost));
    return FX_OK;
  }
};

// time one step of a launch on the stream: GPU ms from events, wall ms added
template <class Launch>
static int timed(Launch launch, hipStream_t st, double* gpu_ms, double* wall_ms) {
  const auto w0 = std::chrono::steady_clock::now();
  hipEvent_t e0, e1;
  FX_HIP(hipEventCreate(&e0));
  FX_HIP(hipEventCreate(&e1));
  hipError_t h = hipEventRecord(e0, st);
  if (h == hipSuccess) {
    launch();
    h = hipGetLastError();
  }
  if (h == hipSuccess) h = hipEventRecord(e1, st);
  if (h == hipSuccess) h = hipEventSynchronize(e1);
  float t = 0.0f;
  if (h == hipSuccess) h = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (h != hipSuccess) return fail_hip(h, "brick advection step");
  *gpu_ms = t;
  *wall_ms += ms_since(w0);
  return FX_OK;
}

// seeded MAC face (mac.mm:76-84)
static float mac_seed(int x, int y, int z, int a, int n, int mode) {
  float p[3] = {x + 0.5f, y + 0.5f, z + 0.5f};
  p[a] -= 0.5f;
  if (mode == 0 || mode == 3) {
    const float lo = n * 0.25f, hi = n * 0.75f;
    for (float v : p)
      if (v < lo || v >= hi) return 0.0f;
  }
  if (mode == 2) return a == 0 ? 0.25f : a == 1 ? -0.125f : 0.0625f;
  if (mode == 3) return a == 0 ? 0.35f : a == 1 ? -0.2f : 0.1f;
  return 0.1f * (a + 1) + 0.2f * (p[(a + 1) % 3] / n - 0.5f) - 0.3f * (p[(a + 2) % 3] / n - 0.5f);
}

}  // namespace ba
}  // namespace fx

using namespace fx;
using namespace fx::ba;

extern "C" {

int fx_transport_compare(uint32_t n, uint32_t side, uint32_t steps, float dx, float dy, float dz, uint64_t budget,
                         float* density, fx_advect_report* r, void* stream) {
  FX_REQUIRE((n == 32 || n == 64 || n == 128 || n == 256) && (side == 8 || side == 16) && steps >= 1 && steps <= 32 &&
                 budget > 0 && budget <= (uint64_t(1) << 30),
             FX_EINVAL, "resolution 32/64/128/256; side 8/16; steps 1..32; budget <=1 GiB per engine");
  const float delta[3] = {dx, dy, dz};
  for (float v : delta)
    FX_REQUIRE(std::isfinite(v) && std::fabs(v) <= 4.0f, FX_EINVAL,
               "Displacement must be finite, at most 4 voxels per step");
  FX_REQUIRE(density && r, FX_EINVAL, "null output");
  hipStream_t st = S(stream);
  *r = fx_advect_report{};
  r->resolution = n;
  r->side = side;
  r->steps = steps;
  const int N = int(n), s = int(side), g = N / s;
  const int extent = N * 7 / 16, lower = (N - extent) / 2;
  auto t0 = std::chrono::steady_clock::now();
  // each step reaches ceil(|d|) cells further along d, and interpolation one
  // cell either way (transport.mm:103-110)
  int lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    const int reach = int(std::ceil(std::fabs(delta[k]))) * int(steps) + 1;
    lo[k] = std::max(0, lower - (delta[k] < 0 ? reach : 1)) / s;
    hi[k] = (std::min(N, lower + extent + (delta[k] > 0 ? reach : 1)) + s - 1) / s;
  }
  std::vector<int> map, coords;
  box_topology(g, lo, hi, map, coords);
  r->active_bricks = uint32_t(coords.size() / 3);
  r->topology_ms = ms_since(t0);
  r->initial_mass = double(extent) * extent * extent;
  Pair sparse(N, s, true, map, coords, false), dense(N, s, false, {}, {}, false);
  FX_REQUIRE(sparse.bytes <= budget && dense.bytes <= budget, FX_EINVAL, "Transport buffers exceed per-engine budget");
  r->sparse_bytes = sparse.bytes;
  r->dense_bytes = dense.bytes;
  auto seed = [&](const Pair& p, std::vector<float>& f) {
    f.assign(size_t(p.L.cells), 0.0f);
    for (size_t i = 0; i < f.size(); ++i) {
      int c[3];
      p.cell(i, c);
      bool in = true;
      for (int k = 0; k < 3; ++k) in = in && c[k] >= lower && c[k] < lower + extent;
      f[i] = in ? 1.0f : 0.0f;
    }
  };
  {
    std::vector<float> init, zero;
    seed(sparse, init);
    zero.assign(init.size(), 0.0f);
    if (int e = sparse.allocate(init, zero, st)) return e;
    seed(dense, init);
    zero.assign(init.size(), 0.0f);
    if (int e = dense.allocate(init, zero, st)) return e;
  }
  // the sample point is cell - delta for every cell: one corner offset and
  // one set of weights
  int base[3];
  float frac[3];
  for (int k = 0; k < 3; ++k) {
    const float o = -delta[k], b = std::floor(o);
    base[k] = int(b);
    frac[k] = o - b;
  }
  auto step = [&](Pair& p) {
    return [&p, &base, &frac, st]() {
      hipLaunchKernelGGL(k_translate, dim3((p.L.cells + 255) / 256), dim3(256), 0, st, p.L, p.f[p.cur], p.f[1 - p.cur],
                         base[0], base[1], base[2], frac[0], frac[1], frac[2]);
      p.cur = 1 - p.cur;
    };
  };
  // alternate the order to spread scheduling bias over both engines
  for (uint32_t i = 0; i < steps; ++i) {
    if (i % 2)
      if (int e = timed(step(dense), st, &r->dense_gpu_ms[i], &r->dense_wall_ms)) return e;
    if (int e = timed(step(sparse), st, &r->sparse_gpu_ms[i], &r->sparse_wall_ms)) return e;
    if (!(i % 2))
      if (int e = timed(step(dense), st, &r->dense_gpu_ms[i], &r->dense_wall_ms)) return e;
  }
  std::vector<float> a, b;
  if (int e = sparse.download(a)) return e;
  if (int e = dense.download(b)) return e;
  double square = 0;
  for (int z = 0; z < N; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < N; ++x) {
        const size_t i = (size_t(z) * N + y) * N + x;
        const long k = sparse.cell_index(x, y, z);
        const float v = k < 0 ? 0.0f : a[size_t(k)];
        FX_REQUIRE(std::isfinite(v) && std::isfinite(b[i]), FX_ESTATE, "Nonfinite transport output");
        const double e = std::fabs(double(v) - b[i]);
        r->max_error = std::max(r->max_error, e);
        square += e * e;
        r->sparse_mass += v;
        r->dense_mass += b[i];
        density[i] = v;
      }
  r->rms_error = std::sqrt(square / (double(N) * N * N));
  return FX_OK;
}

int fx_mac_compare(uint32_t n, uint32_t side, uint32_t steps, uint32_t mode, uint64_t budget, float* faces,
                   fx_advect_report* r, void* stream) {
  FX_REQUIRE((n == 32 || n == 64 || n == 128) && (side == 8 || side == 16) && steps <= 8 && mode <= 3 && budget > 0 &&
                 budget <= (uint64_t(1) << 30),
             FX_EINVAL, "MAC: resolution 32/64/128; side 8/16; steps 0..8; mode 0..3; budget <=1 GiB");
  FX_REQUIRE(faces && r, FX_EINVAL, "null output");
  hipStream_t st = S(stream);
  *r = fx_advect_report{};
  r->resolution = n;
  r->side = side;
  r->steps = steps;
  r->mode = mode;
  const int N = int(n), s = int(side), g = N / s, M = int(mode);
  auto t0 = std::chrono::steady_clock::now();
  // speeds stay below 0.55 voxels per step: two cells per step plus two
  // cover the staggering and the interpolation (mac.mm:93-97)
  const int reach = int(steps) * 2 + 2;
  const bool full = M == 1 || M == 2;
  const int low = full ? 0 : std::max(0, N / 4 - reach) / s;
  const int high = full ? g : (std::min(N, 3 * N / 4 + reach) + s - 1) / s;
  const int lo[3] = {low, low, low}, hi[3] = {high, high, high};
  std::vector<int> map, coords;
  box_topology(g, lo, hi, map, coords);
  r->active_bricks = uint32_t(coords.size() / 3);
  r->topology_ms = ms_since(t0);
  Pair sparse(N, s, true, map, coords, true), dense(N, s, false, {}, {}, true);
  FX_REQUIRE(sparse.bytes <= budget && dense.bytes <= budget, FX_EINVAL, "MAC buffers exceed per-engine budget");
  r->sparse_bytes = sparse.bytes;
  r->dense_bytes = dense.bytes;
  auto seed = [&](const Pair& p, std::vector<float>& init, std::vector<float>& other) {
    const size_t count = size_t(p.L.faces);
    init.assign(count * 3, 0.0f);
    other.assign(count * 3, 0.0f);
    for (int a = 0; a < 3; ++a)
      for (size_t i = 0; i < count; ++i) {
        int c[3];
        const bool owned = p.face(a, i, c);
        init[a * count + i] = owned ? mac_seed(c[0], c[1], c[2], a, N, M) : kPadding;
        other[a * count + i] = owned ? 0.0f : kPadding;
      }
  };
  {
    std::vector<float> init, other;
    seed(sparse, init, other);
    if (int e = sparse.allocate(init, other, st)) return e;
    seed(dense, init, other);
    if (int e = dense.allocate(init, other, st)) return e;
  }
  auto step = [&](Pair& p) {
    return [&p, st]() {
      hipLaunchKernelGGL(k_mac_advect, dim3((p.L.faces + 255) / 256, 3), dim3(256), 0, st, p.L, p.f[p.cur],
                         p.f[1 - p.cur]);
      p.cur = 1 - p.cur;
    };
  };
  for (uint32_t i = 0; i < steps; ++i) {
    if (i % 2)
      if (int e = timed(step(dense), st, &r->dense_gpu_ms[i], &r->dense_wall_ms)) return e;
    if (int e = timed(step(sparse), st, &r->sparse_gpu_ms[i], &r->sparse_wall_ms)) return e;
    if (!(i % 2))
      if (int e = timed(step(dense), st, &r->dense_gpu_ms[i], &r->dense_wall_ms)) return e;
  }
  std::vector<float> a, b;
  if (int e = sparse.download(a)) return e;
  if (int e = dense.download(b)) return e;
  const size_t sc = size_t(sparse.L.faces), dc = size_t(N + 1) * N * N;
  double square = 0;
  for (int ax = 0; ax < 3; ++ax) {
    for (size_t i = 0; i < sc; ++i) {
      int c[3];
      if (sparse.face(ax, i, c)) {
        ++r->owned_faces;
      } else {
        ++r->padding_faces;
        r->padding_error = std::max(r->padding_error, std::fabs(double(a[ax * sc + i]) - kPadding));
      }
    }
    const int sx = N + (ax == 0), sy = N + (ax == 1);
    for (size_t i = 0; i < dc; ++i) {
      const int x = int(i % sx), y = int((i / sx) % sy), z = int(i / (size_t(sx) * sy));
      const long k = sparse.face_index(ax, x, y, z);
      const float v = k < 0 ? 0.0f : a[size_t(k)];
      FX_REQUIRE(std::isfinite(v) && std::isfinite(b[ax * dc + i]), FX_ESTATE, "Nonfinite MAC output");
      const double e = std::fabs(double(v) - b[ax * dc + i]);
      square += e * e;
      r->max_error = std::max(r->max_error, e);
      faces[ax * dc + i] = v;
    }
  }
  r->rms_error = std::sqrt(square / (3.0 * double(dc)));
  return FX_OK;
}

}  // extern "C"
