// Level operators shared by the multigrid engine and the Krylov/IR solvers.
#pragma once
#include "common.hpp"

namespace fx {

// Division of 0 <= n < 2^31 by an invariant d >= 1 as a multiply-high, an add
// and a shift (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, section 4): with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (mulhi(n, m) + n) >> l.  The sum
// cannot wrap since mulhi(n, m) < n < 2^31.  A power of two has m = 1.
struct FastDiv {
  uint32_t m;
  uint32_t l;
};

inline FastDiv make_fastdiv(int d) {
  const uint64_t dd = d > 1 ? uint64_t(d) : 1;
  uint32_t l = 0;
  while ((uint64_t(1) << l) < dd) ++l;
  const uint64_t m = (((uint64_t(1) << l) - dd) << 32) / dd + 1;
  return FastDiv{uint32_t(m), l};
}

__device__ __forceinline__ int fdiv(int n, FastDiv f) {
  return int((__umulhi(uint32_t(n), f.m) + uint32_t(n)) >> f.l);
}

struct LevelDev {
  int nx, ny, nz;
  float wx, wy, wz;
  const void* op;
  // divisors of the vector kernels' lane -> (row, plane) maps: nx/4 and nx/8
  // lanes per row, ny rows per plane, ny/2 coarse rows per plane (1 where the
  // level's extents do not divide)
  FastDiv row4, row8, ny1, ny2;
};

static LevelDev dev_level(const fx_mg_level& L) {
  return LevelDev{L.nx, L.ny, L.nz, L.w[0], L.w[1], L.w[2], L.op_data,
                  make_fastdiv(L.nx / 4), make_fastdiv(L.nx / 8), make_fastdiv(L.ny), make_fastdiv(L.ny / 2)};
}

// lanes-per-row divisor of a W-cell-per-lane kernel
template <int W>
__device__ __forceinline__ FastDiv row_div(const LevelDev& L) {
  static_assert(W == 4 || W == 8, "vector kernels hold 4 or 8 cells per lane");
  return W == 4 ? L.row4 : L.row8;
}

struct Stencil {
  float s;       // sum_n w_n (p_n - p_c): neighbour differences first, so an
                 // offset in p does not cancel catastrophically in fp32
  float d;       // sum_n w_n
  float c;       // p_c
  bool active;   // fluid / active cell
};

constexpr bool op_has_edges(int op) { return op == FX_OP_EDGES || op == FX_OP_PACKED; }

// (w to +x, w to +y, w to +z, active) of cell i of an EDGES or PACKED level
template <int OP>
__device__ __forceinline__ float4 edge_at(const LevelDev& L, int64_t i) {
  if (OP == FX_OP_EDGES) return reinterpret_cast<const float4*>(L.op)[i];
  const uint32_t u = reinterpret_cast<const uint32_t*>(L.op)[i];
  return make_float4(float(u & 1023u) * L.wx, float((u >> 10) & 1023u) * L.wy, float((u >> 20) & 1023u) * L.wz,
                     (u >> 30) & 1u ? 1.0f : 0.0f);
}
template <int OP>
__device__ __forceinline__ float edge_x(const LevelDev& L, int64_t i) {
  if (OP == FX_OP_EDGES) return reinterpret_cast<const float4*>(L.op)[i].x;
  return float(reinterpret_cast<const uint32_t*>(L.op)[i] & 1023u) * L.wx;
}
template <int OP>
__device__ __forceinline__ float edge_y(const LevelDev& L, int64_t i) {
  if (OP == FX_OP_EDGES) return reinterpret_cast<const float4*>(L.op)[i].y;
  return float((reinterpret_cast<const uint32_t*>(L.op)[i] >> 10) & 1023u) * L.wy;
}
template <int OP>
__device__ __forceinline__ float edge_z(const LevelDev& L, int64_t i) {
  if (OP == FX_OP_EDGES) return reinterpret_cast<const float4*>(L.op)[i].z;
  return float((reinterpret_cast<const uint32_t*>(L.op)[i] >> 20) & 1023u) * L.wz;
}

// Gather the 7-point operator at cell i for the given level operator.
// read_p = false: p is known to be zero (s = c = 0).
template <int OP, typename T>
__device__ __forceinline__ Stencil stencil(const LevelDev& L, int x, int y, int z, int64_t i,
                                           const T* __restrict__ p, bool read_p) {
  const int64_t sy = L.nx, sz = int64_t(L.nx) * L.ny;
  Stencil st{0.0f, 0.0f, 0.0f, true};
  const float c0 = read_p ? to_f(p[i]) : 0.0f;
  st.c = c0;
  if (op_has_edges(OP)) {
    float4 c = edge_at<OP>(L, i);
    float wxl = x > 0 ? edge_x<OP>(L, i - 1) : 0.0f;
    float wyl = y > 0 ? edge_y<OP>(L, i - sy) : 0.0f;
    float wzl = z > 0 ? edge_z<OP>(L, i - sz) : 0.0f;
    st.d = c.x + c.y + c.z + wxl + wyl + wzl;
    st.active = c.w > 0.0f;
    if (read_p) {
      float s = 0.0f;
      if (c.x != 0.0f) s += c.x * (to_f(p[i + 1]) - c0);
      if (wxl != 0.0f) s += wxl * (to_f(p[i - 1]) - c0);
      if (c.y != 0.0f) s += c.y * (to_f(p[i + sy]) - c0);
      if (wyl != 0.0f) s += wyl * (to_f(p[i - sy]) - c0);
      if (c.z != 0.0f) s += c.z * (to_f(p[i + sz]) - c0);
      if (wzl != 0.0f) s += wzl * (to_f(p[i - sz]) - c0);
      st.s = s;
    }
  } else {
    uint8_t f = OP == FX_OP_FLAGS ? reinterpret_cast<const uint8_t*>(L.op)[i] : box_flags(x, y, z, L.nx, L.ny, L.nz);
    st.active = f & FLAG_FLUID;
    float n[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (read_p) {
      if (f & FLAG_XLO) n[0] = to_f(p[i - 1]);
      if (f & FLAG_XHI) n[1] = to_f(p[i + 1]);
      if (f & FLAG_YLO) n[2] = to_f(p[i - sy]);
      if (f & FLAG_YHI) n[3] = to_f(p[i + sy]);
      if (f & FLAG_ZLO) n[4] = to_f(p[i - sz]);
      if (f & FLAG_ZHI) n[5] = to_f(p[i + sz]);
    }
    flag_cell(f, c0, n[0], n[1], n[2], n[3], n[4], n[5], L.wx, L.wy, L.wz, st.s, st.d);
  }
  return st;
}

template <int OP>
__device__ __forceinline__ bool cell_active(const LevelDev& L, int x, int y, int z, int64_t i) {
  if (OP == FX_OP_EDGES) return reinterpret_cast<const float4*>(L.op)[i].w > 0.0f;
  if (OP == FX_OP_PACKED) return (reinterpret_cast<const uint32_t*>(L.op)[i] >> 30) & 1u;
  if (OP == FX_OP_FLAGS) return reinterpret_cast<const uint8_t*>(L.op)[i] & FLAG_FLUID;
  return true;
}


// residual r = rhs*inv_dt - (s - d p) on active cells, 0 elsewhere
template <typename T, int OP>
__device__ __forceinline__ float residual_at(const LevelDev& L, int x, int y, int z, const T* __restrict__ p,
                                             const T* __restrict__ rhs, float inv_dt) {
  int64_t i = idx3(x, y, z, L.nx, L.ny);
  Stencil st = stencil<OP>(L, x, y, z, i, p, true);
  if (!st.active) return 0.0f;
  if (op_has_edges(OP) && !(st.d > 0.0f)) return 0.0f;
  return to_f(rhs[i]) * inv_dt - st.s;
}

static const dim3 kMgBlock(64, 4, 1);
inline dim3 level_grid(const fx_mg_level& L) { return grid3(L.nx, L.ny, L.nz, kMgBlock); }

#define FX_DISPATCH_OP(op, ...)                                             \
  switch (op) {                                                             \
    case FX_OP_BOX: { constexpr int OPV = FX_OP_BOX; __VA_ARGS__; break; }  \
    case FX_OP_FLAGS: { constexpr int OPV = FX_OP_FLAGS; __VA_ARGS__; break; } \
    case FX_OP_EDGES: { constexpr int OPV = FX_OP_EDGES; __VA_ARGS__; break; } \
    case FX_OP_PACKED: { constexpr int OPV = FX_OP_PACKED; __VA_ARGS__; break; } \
    default: set_error("unknown level operator %d", op); return FX_EINVAL; \
  }

#define FX_DISPATCH_T(st, ...)                                              \
  switch (st) {                                                             \
    case FX_F32: { using TV = float; __VA_ARGS__; break; }                  \
    case FX_BF16: { using TV = bf16; __VA_ARGS__; break; }                  \
    case FX_F16: { using TV = _Float16; __VA_ARGS__; break; }               \
    default: set_error("unknown storage type %d", st); return FX_EINVAL;  \
  }

// host-side engine pieces (multigrid.hip)
int check_config(const fx_mg_config* c);
int vcycle(fx_mg_config* c, int l, float inv_dt, bool zero_guess, hipStream_t s);
int resnorm(const fx_mg_level& L0, const float* x, const float* b, double* ws, double* rn, double* bn,
            hipStream_t s);
int project_mean_level(const fx_mg_level& L0, float* b, double* ws, hipStream_t s);
fx_mg_level level0_of(fx_grid g, const uint8_t* flags);

}  // namespace fx

// Cell loop of the 1-D reduction grids: one wave per (y, z) row with lanes
// along x, so cell coordinates need no per-cell integer division.
#define FX_FOR_ROW_CELLS(L, X, Y, Z, i)                                                              \
  for (int row_ = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), rows_ = (L).ny * (L).nz;       \
       row_ < rows_; row_ += gridDim.x * (blockDim.x >> 6))                                          \
    for (int Z = row_ / (L).ny, Y = row_ - Z * (L).ny, X = threadIdx.x & 63; X < (L).nx; X += 64)     \
      if (const int64_t i = int64_t(row_) * (L).nx + X; true)
