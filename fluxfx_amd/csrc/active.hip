// Exact active-brick sets of the scalar passes.  A scalar channel is
// described per 8^3 brick by a live byte: 0 guarantees every cell of the brick
// holds exactly 0 (1 allows anything).  The transport of all-zero scalars is
// exactly zero wherever neither a nonzero value nor an emitter lies within the
// reach of its traces, so the scalar predictor, the scalar transport and the
// combustion pass run only on the bricks a plan selects, and give the same
// values as over the whole grid:
//
//   transport  N = live inputs dilated by one brick, plus emitter bricks:
//              computed; outputs live elsewhere: stored 0; the rest skipped
//              P = N dilated by one brick (where the transport reads
//              predictors): the predictor computes these
//   combustion N = live inputs (pointwise: a cell without fuel, heat or smoke
//              stays zero): computed; outputs live elsewhere: stored 0
//
// Transport plans hold while every trace stays within 2.5 cells of its site:
// semi-Lagrangian donors then lie within 3 cells and the MacCormack reverse
// sample reads predictors within 3 cells whose donors lie within 3 more, so
// 6 < 8 cells.  A plan whose velocity bound dt * max|u| / h exceeds 2.5 (or is
// not finite) selects every brick.
#include "common.hpp"

namespace fx {

constexpr int kMaxPlanChannels = 4;
constexpr int kMaxEmitterBoxes = 8;

__global__ void k_active_zero(int64_t n, uint8_t* __restrict__ a) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    a[i] = 0;
}

// live[b] = 1 for every brick holding a value that is not exactly 0 (NaN
// included); the mask has been cleared.  Row form: nx % 4 == 0 and a 16-byte
// aligned field, one float4 per thread.
__global__ void k_active_scan4(int nx, int ny, int bx, int by, int64_t nq, const float4* __restrict__ f,
                               uint8_t* __restrict__ live) {
  const int qx = nx >> 2;
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nq; q += int64_t(gridDim.x) * blockDim.x) {
    const float4 v = f[q];
    if (!(v.x == 0.0f) || !(v.y == 0.0f) || !(v.z == 0.0f) || !(v.w == 0.0f)) {
      const int64_t row = q / qx;
      const int x = int(q - row * qx) << 2, y = int(row % ny), z = int(row / ny);
      live[(x >> kActiveShift) + int64_t(bx) * ((y >> kActiveShift) + int64_t(by) * (z >> kActiveShift))] = 1;
    }
  }
}

__global__ void k_active_scan(int nx, int ny, int bx, int by, int64_t n, const float* __restrict__ f,
                              uint8_t* __restrict__ live) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    if (!(f[i] == 0.0f)) {
      const int64_t row = i / nx;
      const int x = int(i - row * nx), y = int(row % ny), z = int(row / ny);
      live[(x >> kActiveShift) + int64_t(bx) * ((y >> kActiveShift) + int64_t(by) * (z >> kActiveShift))] = 1;
    }
  }
}

struct PlanArgs {
  int bx, by, bz;
  int transport;  // 1: transport plan (dilations, emitters, velocity bound); 0: pointwise
  int nin, nout;
  const uint8_t* in[kMaxPlanChannels];
  uint8_t* out[kMaxPlanChannels];
  uint8_t* predict;  // P (transport plans)
  uint8_t* write;    // W
  const float* vmax;  // device max |U|, |V|, |W| (transport plans)
  float sx, sy, sz;   // dt * cells per metre along each axis
  int nbox;
  int box[kMaxEmitterBoxes][6];  // emitter bricks, inclusive [x0, x1] x [y0, y1] x [z0, z1]
};

// A block owns a 4 x 4 x 4 tile of bricks, one thread per brick, and stages
// the union of the input masks over the tile and two bricks around it in LDS
// (0 outside the grid).  The thread also clears the live bytes of the output
// buffers of every brick the pass writes (the pass sets them again where it
// stores nonzero values), so each output's mask stays exact: a skipped brick
// keeps its zeros, a written one is rewritten.
constexpr int kPlanTile = 4, kPlanHalo = 2, kPlanSpan = kPlanTile + 2 * kPlanHalo;

__global__ __launch_bounds__(kPlanTile * kPlanTile * kPlanTile) void k_active_plan(PlanArgs A) {
  __shared__ uint8_t U[kPlanSpan][kPlanSpan][kPlanSpan];
  const int t = threadIdx.x;
  const int ox = blockIdx.x * kPlanTile, oy = blockIdx.y * kPlanTile, oz = blockIdx.z * kPlanTile;
  for (int k = t; k < kPlanSpan * kPlanSpan * kPlanSpan; k += blockDim.x) {
    const int hx = k % kPlanSpan, hy = (k / kPlanSpan) % kPlanSpan, hz = k / (kPlanSpan * kPlanSpan);
    const int x = ox + hx - kPlanHalo, y = oy + hy - kPlanHalo, z = oz + hz - kPlanHalo;
    uint8_t any = 0;
    if (x >= 0 && y >= 0 && z >= 0 && x < A.bx && y < A.by && z < A.bz) {
      const int b = x + A.bx * (y + A.by * z);
      for (int c = 0; c < A.nin; ++c) any |= A.in[c][b];
    }
    U[hz][hy][hx] = any;
  }
  __syncthreads();
  const int tx = t % kPlanTile, ty = (t / kPlanTile) % kPlanTile, tz = t / (kPlanTile * kPlanTile);
  const int cx = ox + tx, cy = oy + ty, cz = oz + tz;
  if (cx >= A.bx || cy >= A.by || cz >= A.bz) return;
  const int b = cx + A.bx * (cy + A.by * cz);
  auto live = [&](int dx, int dy, int dz) {
    return U[tz + kPlanHalo + dz][ty + kPlanHalo + dy][tx + kPlanHalo + dx];
  };
  bool n = false, p = false;
  if (A.transport) {
    const float d = fmaxf(fmaxf(A.vmax[0] * A.sx, A.vmax[1] * A.sy), A.vmax[2] * A.sz);
    if (!(d <= 2.5f)) {
      n = p = true;
    } else {
      // all 125 reads issued together (no early exit): OR within one and two bricks
      uint32_t near = 0, far = 0;
#pragma unroll
      for (int dz = -2; dz <= 2; ++dz)
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy)
#pragma unroll
          for (int dx = -2; dx <= 2; ++dx) {
            const uint32_t v = live(dx, dy, dz);
            far |= v;
            if (abs(dx) <= 1 && abs(dy) <= 1 && abs(dz) <= 1) near |= v;
          }
      n = near != 0;
      p = far != 0;
      for (int s = 0; s < A.nbox; ++s) {
        const int* e = A.box[s];
        const int ox = cx < e[0] ? e[0] - cx : (cx > e[1] ? cx - e[1] : 0);
        const int oy = cy < e[2] ? e[2] - cy : (cy > e[3] ? cy - e[3] : 0);
        const int oz = cz < e[4] ? e[4] - cz : (cz > e[5] ? cz - e[5] : 0);
        const int o = max(ox, max(oy, oz));
        n = n || o == 0;
        p = p || o <= 1;
      }
    }
  } else {
    n = live(0, 0, 0) != 0;
  }
  bool stale = false;
  for (int c = 0; c < A.nout; ++c) stale = stale || A.out[c][b] != 0;
  if (A.predict) A.predict[b] = p ? kBrickCompute : kBrickSkip;
  A.write[b] = n ? kBrickCompute : (stale ? kBrickZero : kBrickSkip);
  if (n || stale)
    for (int c = 0; c < A.nout; ++c) A.out[c][b] = 0;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_active_scan(fx_grid g, const float* field, uint8_t* live, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(field && live, FX_EINVAL, "active scan: null array");
  const int bx = active_bricks_along(g.nx), by = active_bricks_along(g.ny), bz = active_bricks_along(g.nz);
  const int64_t nb = int64_t(bx) * by * bz, n = int64_t(g.nx) * g.ny * g.nz;
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_active_zero, dim3(unsigned(std::min<int64_t>((nb + 255) / 256, 1024))), dim3(256), 0, s, nb,
                     live);
  if (g.nx % 4 == 0 && reinterpret_cast<uintptr_t>(field) % 16 == 0) {
    const int64_t nq = n / 4;
    hipLaunchKernelGGL(k_active_scan4, dim3(unsigned(std::min<int64_t>((nq + 255) / 256, 8192))), dim3(256), 0, s,
                       g.nx, g.ny, bx, by, nq, reinterpret_cast<const float4*>(field), live);
  } else {
    hipLaunchKernelGGL(k_active_scan, dim3(unsigned(std::min<int64_t>((n + 255) / 256, 8192))), dim3(256), 0, s,
                       g.nx, g.ny, bx, by, n, field, live);
  }
  FX_LAUNCHED("fx_active_scan");
  return FX_OK;
}

int fx_active_plan(fx_grid g, float dt, const float* vmax, int32_t transport, int32_t nin,
                   const uint8_t* const* live_in, int32_t nout, uint8_t* const* live_out, const fx_sources* sources,
                   uint8_t* predict_bricks, uint8_t* write_bricks, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(transport == 0 || transport == 1, FX_EINVAL, "active plan: transport must be 0 or 1");
  FX_REQUIRE(nin >= 1 && nin <= kMaxPlanChannels && nout >= 0 && nout <= kMaxPlanChannels && live_in && write_bricks &&
                 (nout == 0 || live_out),
             FX_EINVAL, "active plan: 1..%d input and 0..%d output masks", kMaxPlanChannels, kMaxPlanChannels);
  PlanArgs A{};
  A.bx = active_bricks_along(g.nx);
  A.by = active_bricks_along(g.ny);
  A.bz = active_bricks_along(g.nz);
  A.transport = transport;
  A.nin = nin;
  A.nout = nout;
  for (int c = 0; c < nin; ++c) {
    FX_REQUIRE(live_in[c], FX_EINVAL, "active plan: null input mask %d", c);
    A.in[c] = live_in[c];
  }
  for (int c = 0; c < nout; ++c) {
    FX_REQUIRE(live_out[c] && live_out[c] != write_bricks && live_out[c] != predict_bricks, FX_EINVAL,
               "active plan: output mask %d must be distinct from the plan", c);
    for (int j = 0; j < nin; ++j)
      FX_REQUIRE(live_out[c] != live_in[j], FX_EINVAL, "active plan: output mask %d is an input mask", c);
    A.out[c] = live_out[c];
  }
  A.write = write_bricks;
  if (transport) {
    FX_REQUIRE(vmax && predict_bricks && predict_bricks != write_bricks && sources, FX_EINVAL,
               "active plan: a transport plan needs the velocity bound, both plan masks and the sources");
    FX_REQUIRE(dt > 0.0f && std::isfinite(dt), FX_EINVAL, "active plan: dt must be positive");
    FX_REQUIRE(sources->count >= 0 && sources->count <= kMaxEmitterBoxes, FX_EINVAL,
               "active plan: at most %d sources", kMaxEmitterBoxes);
    A.predict = predict_bricks;
    A.vmax = vmax;
    A.sx = float(double(dt) * g.nx / g.ex);
    A.sy = float(double(dt) * g.ny / g.ey);
    A.sz = float(double(dt) * g.nz / g.ez);
    // an emitter weighs cells within two radii of the box around its swept
    // segment (centre and start): the cells whose centres lie in that box, one
    // cell wider, give the bricks it may write into
    const int n3[3] = {g.nx, g.ny, g.nz};
    const double h3[3] = {double(g.ex) / g.nx, double(g.ey) / g.ny, double(g.ez) / g.nz};
    for (int k = 0; k < sources->count; ++k) {
      const float* row = sources->table + 16 * k;
      bool inside = true;
      int e[6];
      for (int a = 0; a < 3; ++a) {
        FX_REQUIRE(std::isfinite(row[a]) && std::isfinite(row[4 + a]) && std::isfinite(row[3]), FX_EINVAL,
                   "active plan: source %d has a non-finite position", k);
        const double lo = std::min(row[a], row[4 + a]) - 2.0 * row[3], hi = std::max(row[a], row[4 + a]) + 2.0 * row[3];
        const double c0 = std::floor(lo / h3[a] - 0.5) - 1.0, c1 = std::ceil(hi / h3[a] - 0.5) + 1.0;
        if (c1 < 0.0 || c0 > n3[a] - 1) inside = false;
        const int i0 = int(std::max(c0, 0.0)), i1 = int(std::min(c1, double(n3[a] - 1)));
        e[2 * a] = i0 >> kActiveShift;
        e[2 * a + 1] = i1 >> kActiveShift;
      }
      if (!inside) continue;
      for (int j = 0; j < 6; ++j) A.box[A.nbox][j] = e[j];
      ++A.nbox;
    }
  }
  const dim3 grid((A.bx + kPlanTile - 1) / kPlanTile, (A.by + kPlanTile - 1) / kPlanTile,
                  (A.bz + kPlanTile - 1) / kPlanTile);
  hipLaunchKernelGGL(k_active_plan, grid, dim3(kPlanTile * kPlanTile * kPlanTile), 0, S(stream), A);
  FX_LAUNCHED("fx_active_plan");
  return FX_OK;
}

}  // extern "C"
