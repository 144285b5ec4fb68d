// Four-cells-per-lane ("quad") access to x-fast level arrays.
//
// A lane owns four consecutive x cells (x0 % 4 == 0) of one row, so every
// array is read with one 16-byte (fp32) or 8-byte (bf16/fp16) vector load per
// quad instead of four scalar loads.  The x neighbours outside the quad come
// from two scalar loads; y and z neighbours are quad loads of the adjacent
// rows/planes (cache hits for the neighbouring lanes' data).  Usable when
// nx % 4 == 0 and every array is 16-byte aligned; the 7-point sums keep the
// order of the scalar stencil (x-, x+, y-, y+, z-, z+), so results match the
// one-cell kernels bit for bit.
#pragma once
#include "mg_internal.hpp"

namespace fx {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

template <int W>
struct FV {
  float v[W];
};
using F4 = FV<4>;

// cells per lane for a 16-byte vector of storage type T
template <typename T>
constexpr int vec_width() {
  return int(16 / sizeof(T));
}

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

template <typename T>
__device__ __forceinline__ F4 ld4(const T* __restrict__ p, int64_t i);
template <>
__device__ __forceinline__ F4 ld4<float>(const float* __restrict__ p, int64_t i) {
  float4 a = *reinterpret_cast<const float4*>(p + i);
  return F4{{a.x, a.y, a.z, a.w}};
}
template <>
__device__ __forceinline__ F4 ld4<bf16>(const bf16* __restrict__ p, int64_t i) {
  uint2 a = *reinterpret_cast<const uint2*>(p + i);
  return F4{{__uint_as_float(a.x << 16), __uint_as_float(a.x & 0xffff0000u), __uint_as_float(a.y << 16),
             __uint_as_float(a.y & 0xffff0000u)}};
}
template <>
__device__ __forceinline__ F4 ld4<_Float16>(const _Float16* __restrict__ p, int64_t i) {
  half4_t a = *reinterpret_cast<const half4_t*>(p + i);
  return F4{{float(a.x), float(a.y), float(a.z), float(a.w)}};
}

template <typename T>
__device__ __forceinline__ void st4(T* __restrict__ p, int64_t i, const F4& v);
template <>
__device__ __forceinline__ void st4<float>(float* __restrict__ p, int64_t i, const F4& v) {
  *reinterpret_cast<float4*>(p + i) = make_float4(v.v[0], v.v[1], v.v[2], v.v[3]);
}
template <>
__device__ __forceinline__ void st4<bf16>(bf16* __restrict__ p, int64_t i, const F4& v) {
  uint2 a;
  a.x = uint32_t(from_f<bf16>(v.v[0]).bits) | (uint32_t(from_f<bf16>(v.v[1]).bits) << 16);
  a.y = uint32_t(from_f<bf16>(v.v[2]).bits) | (uint32_t(from_f<bf16>(v.v[3]).bits) << 16);
  *reinterpret_cast<uint2*>(p + i) = a;
}
template <>
__device__ __forceinline__ void st4<_Float16>(_Float16* __restrict__ p, int64_t i, const F4& v) {
  half4_t a = {from_f<_Float16>(v.v[0]), from_f<_Float16>(v.v[1]), from_f<_Float16>(v.v[2]),
               from_f<_Float16>(v.v[3])};
  *reinterpret_cast<half4_t*>(p + i) = a;
}

// generic W-wide loads/stores (W * sizeof(T) is 8 or 16 bytes)
template <typename T, int W>
struct VecIO;
template <>
struct VecIO<float, 4> {
  static __device__ __forceinline__ FV<4> ld(const float* __restrict__ p, int64_t i) { return ld4(p, i); }
  static __device__ __forceinline__ void st(float* __restrict__ p, int64_t i, const FV<4>& v) { st4(p, i, v); }
};
template <>
struct VecIO<bf16, 4> {
  static __device__ __forceinline__ FV<4> ld(const bf16* __restrict__ p, int64_t i) { return ld4(p, i); }
  static __device__ __forceinline__ void st(bf16* __restrict__ p, int64_t i, const FV<4>& v) { st4(p, i, v); }
};
template <>
struct VecIO<_Float16, 4> {
  static __device__ __forceinline__ FV<4> ld(const _Float16* __restrict__ p, int64_t i) { return ld4(p, i); }
  static __device__ __forceinline__ void st(_Float16* __restrict__ p, int64_t i, const FV<4>& v) { st4(p, i, v); }
};
template <>
struct VecIO<bf16, 8> {
  static __device__ __forceinline__ FV<8> ld(const bf16* __restrict__ p, int64_t i) {
    const uint4 a = *reinterpret_cast<const uint4*>(p + i);
    const uint32_t u[4] = {a.x, a.y, a.z, a.w};
    FV<8> r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      r.v[2 * k] = __uint_as_float(u[k] << 16);
      r.v[2 * k + 1] = __uint_as_float(u[k] & 0xffff0000u);
    }
    return r;
  }
  static __device__ __forceinline__ void st(bf16* __restrict__ p, int64_t i, const FV<8>& v) {
    uint32_t u[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      u[k] = uint32_t(from_f<bf16>(v.v[2 * k]).bits) | (uint32_t(from_f<bf16>(v.v[2 * k + 1]).bits) << 16);
    *reinterpret_cast<uint4*>(p + i) = make_uint4(u[0], u[1], u[2], u[3]);
  }
};
template <>
struct VecIO<_Float16, 8> {
  static __device__ __forceinline__ FV<8> ld(const _Float16* __restrict__ p, int64_t i) {
    const half8_t a = *reinterpret_cast<const half8_t*>(p + i);
    FV<8> r;
#pragma unroll
    for (int k = 0; k < 8; ++k) r.v[k] = float(a[k]);
    return r;
  }
  static __device__ __forceinline__ void st(_Float16* __restrict__ p, int64_t i, const FV<8>& v) {
    half8_t a;
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = from_f<_Float16>(v.v[k]);
    *reinterpret_cast<half8_t*>(p + i) = a;
  }
};

// two consecutive values at an even index
template <typename T>
__device__ __forceinline__ void st2(T* __restrict__ p, int64_t i, float a, float b);
template <>
__device__ __forceinline__ void st2<float>(float* __restrict__ p, int64_t i, float a, float b) {
  *reinterpret_cast<float2*>(p + i) = make_float2(a, b);
}
template <>
__device__ __forceinline__ void st2<bf16>(bf16* __restrict__ p, int64_t i, float a, float b) {
  *reinterpret_cast<uint32_t*>(p + i) = uint32_t(from_f<bf16>(a).bits) | (uint32_t(from_f<bf16>(b).bits) << 16);
}
template <>
__device__ __forceinline__ void st2<_Float16>(_Float16* __restrict__ p, int64_t i, float a, float b) {
  half2_t h = {from_f<_Float16>(a), from_f<_Float16>(b)};
  *reinterpret_cast<half2_t*>(p + i) = h;
}

// `n` (2 or 4) consecutive coarse values at index i (i % n == 0)
template <typename T, int N>
__device__ __forceinline__ void st_coarse(T* __restrict__ p, int64_t i, const float* v) {
  if (N == 2) {
    st2(p, i, v[0], v[1]);
  } else {
    FV<4> f{{v[0], v[1], v[N > 2 ? 2 : 0], v[N > 3 ? 3 : 0]}};
    st4(p, i, f);
  }
}

// Reads of the operand p in storage T.
template <typename T>
struct PLoad {
  const T* __restrict__ p;
  __device__ __forceinline__ F4 quad(int64_t j) const { return ld4(p, j); }
  template <int W>
  __device__ __forceinline__ FV<W> vec(int64_t j) const { return VecIO<T, W>::ld(p, j); }
  __device__ __forceinline__ float one(int64_t j) const { return to_f(p[j]); }
};

// Reads of x + s*e (x fp32, e in storage T or null).
template <typename T>
struct ShiftLoad {
  const float* __restrict__ x;
  const T* __restrict__ e;
  float s;
  __device__ __forceinline__ F4 quad(int64_t j) const {
    F4 a = ld4(x, j);
    if (e) {
      F4 b = ld4(e, j);
#pragma unroll
      for (int k = 0; k < 4; ++k) a.v[k] += s * b.v[k];
    }
    return a;
  }
  __device__ __forceinline__ float one(int64_t j) const { return x[j] + (e ? s * to_f(e[j]) : 0.0f); }
};

struct QuadStencil {
  float s[4];     // sum_n w_n (p_n - p_c)
  float d[4];     // sum_n w_n
  float c[4];     // p_c
  uint32_t f[4];  // open-face / fluid bits
};

// Quad coordinates of quad q of a level (x0 = 4*(q mod nx/4)).
struct QuadPos {
  int x, y, z;
  int64_t i;
};

__device__ __forceinline__ QuadPos quad_pos(const LevelDev& L, int q) {
  const int qx = L.nx >> 2;
  const int row = fdiv(q, L.row4);
  QuadPos r;
  r.x = (q - row * qx) << 2;
  r.z = fdiv(row, L.ny1);
  r.y = row - r.z * L.ny;
  r.i = int64_t(row) * L.nx + r.x;
  return r;
}

__device__ __forceinline__ uint32_t quad_flags(const LevelDev& L, int op, int x0, int y, int z, int64_t i) {
  if (op == FX_OP_FLAGS) return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint8_t*>(L.op) + i);
  uint32_t w = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) w |= uint32_t(box_flags(x0 + k, y, z, L.nx, L.ny, L.nz)) << (8 * k);
  return w;
}

// 7-point operator over the quad at (x0, y, z) for a FLAGS or BOX level.
template <int OP, class Ld>
__device__ __forceinline__ QuadStencil quad_stencil(const LevelDev& L, int x0, int y, int z, int64_t i, const Ld& ld,
                                                    bool read_p) {
  static_assert(!op_has_edges(OP), "quad stencils cover flag and box levels");
  const int64_t sy = L.nx, sz = int64_t(L.nx) * L.ny;
  const uint32_t fw = quad_flags(L, OP, x0, y, z, i);
  F4 c{}, ym{}, yp{}, zm{}, zp{};
  float xl = 0.0f, xr = 0.0f;
  if (read_p) {
    c = ld.quad(i);
    if (y > 0) ym = ld.quad(i - sy);
    if (y + 1 < L.ny) yp = ld.quad(i + sy);
    if (z > 0) zm = ld.quad(i - sz);
    if (z + 1 < L.nz) zp = ld.quad(i + sz);
    if (x0 > 0) xl = ld.one(i - 1);
    if (x0 + 4 < L.nx) xr = ld.one(i + 4);
  }
  QuadStencil st;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t f = (fw >> (8 * k)) & 0xffu;
    const float ck = c.v[k];
    const float left = k > 0 ? c.v[k > 0 ? k - 1 : 0] : xl;
    const float right = k < 3 ? c.v[k < 3 ? k + 1 : 3] : xr;
    flag_cell(f, ck, left, right, ym.v[k], yp.v[k], zm.v[k], zp.v[k], L.wx, L.wy, L.wz, st.s[k], st.d[k]);
    st.c[k] = ck;
    st.f[k] = f;
  }
  return st;
}

// levels below this size keep the one-cell kernels (more lanes in flight)
constexpr int64_t kQuadMinCells = int64_t(1) << 15;

// W-wide generalisation of QuadStencil / quad_stencil (W = 4 or 8 cells per lane)
template <int W>
struct VecStencil {
  float s[W];
  float d[W];
  float c[W];
  uint32_t f[W];
};

// W consecutive packed operator words at i (i % 4 == 0, 16-byte aligned)
template <int W>
struct UV {
  uint32_t v[W];
};
template <int W>
__device__ __forceinline__ UV<W> ld_words(const uint32_t* __restrict__ p, int64_t i) {
  UV<W> r;
#pragma unroll
  for (int h = 0; h < W / 4; ++h) {
    const uint4 a = *reinterpret_cast<const uint4*>(p + i + 4 * h);
    r.v[4 * h] = a.x;
    r.v[4 * h + 1] = a.y;
    r.v[4 * h + 2] = a.z;
    r.v[4 * h + 3] = a.w;
  }
  return r;
}

// Neighbour operand loads of a W-wide stencil.  The 8-wide (16-bit storage)
// kernels issue every load unconditionally (a neighbour outside the level
// reads the lane's own cells, which a closed face or zero conductance never
// uses), so one wait covers them all and the per-cell branches go; the 4-wide
// fp32 kernels keep the guarded loads, which hold fewer registers (more waves
// in flight).
template <int W, class Ld>
__device__ __forceinline__ void vec_neighbours(const LevelDev& L, int x0, int y, int z, int64_t i, const Ld& ld,
                                               FV<W>& c, FV<W>& ym, FV<W>& yp, FV<W>& zm, FV<W>& zp, float& xl,
                                               float& xr) {
  const int64_t sy = L.nx, sz = int64_t(L.nx) * L.ny;
  c = ld.template vec<W>(i);
  if constexpr (W == 8) {
    ym = ld.template vec<W>(y > 0 ? i - sy : i);
    yp = ld.template vec<W>(y + 1 < L.ny ? i + sy : i);
    zm = ld.template vec<W>(z > 0 ? i - sz : i);
    zp = ld.template vec<W>(z + 1 < L.nz ? i + sz : i);
    xl = ld.one(x0 > 0 ? i - 1 : i);
    xr = ld.one(x0 + W < L.nx ? i + W : i);
  } else {
    if (y > 0) ym = ld.template vec<W>(i - sy);
    if (y + 1 < L.ny) yp = ld.template vec<W>(i + sy);
    if (z > 0) zm = ld.template vec<W>(i - sz);
    if (z + 1 < L.nz) zp = ld.template vec<W>(i + sz);
    if (x0 > 0) xl = ld.one(i - 1);
    if (x0 + W < L.nx) xr = ld.one(i + W);
  }
}

// per-cell bits of W cells: open-face/fluid flags (FLAGS, BOX) or FLAG_FLUID
// for an active cell (PACKED)
template <int W>
__device__ __forceinline__ void vec_flags(const LevelDev& L, int op, int x0, int y, int z, int64_t i, uint32_t* f) {
  if (op == FX_OP_PACKED) {
    const UV<W> w = ld_words<W>(reinterpret_cast<const uint32_t*>(L.op), i);
#pragma unroll
    for (int k = 0; k < W; ++k) f[k] = ((w.v[k] >> 30) & 1u) ? uint32_t(FLAG_FLUID) : 0u;
  } else if (op == FX_OP_FLAGS) {
    const uint8_t* fl = reinterpret_cast<const uint8_t*>(L.op) + i;
    if (W == 4) {
      const uint32_t w = *reinterpret_cast<const uint32_t*>(fl);
#pragma unroll
      for (int k = 0; k < W; ++k) f[k] = (w >> (8 * (k & 3))) & 0xffu;
    } else {
      const uint2 w = *reinterpret_cast<const uint2*>(fl);
#pragma unroll
      for (int k = 0; k < W; ++k) f[k] = ((k < 4 ? w.x : w.y) >> (8 * (k & 3))) & 0xffu;
    }
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) f[k] = box_flags(x0 + k, y, z, L.nx, L.ny, L.nz);
  }
}

// W-wide 7-point operator of a PACKED level: the cell's own word holds the
// conductances to +x/+y/+z, the -x/-y/-z ones come from the words of the cells
// below.  Sums keep the order of the one-cell edge stencil (d = +x +y +z -x -y
// -z; s over +x, -x, +y, -y, +z, -z, zero conductances skipped), so results
// match it bit for bit.  f = FLAG_FLUID for active cells.
template <int W, class Ld>
__device__ __forceinline__ VecStencil<W> vec_stencil_packed(const LevelDev& L, int x0, int y, int z, int64_t i,
                                                            const Ld& ld, bool read_p) {
  const int64_t sy = L.nx, sz = int64_t(L.nx) * L.ny;
  const uint32_t* __restrict__ pk = reinterpret_cast<const uint32_t*>(L.op);
  // the words of missing lower neighbours read as zero (unconditional loads
  // and a select in the 8-wide kernels, guarded loads in the 4-wide ones)
  const UV<W> wo = ld_words<W>(pk, i);
  UV<W> wym{}, wzm{};
  uint32_t wxl = 0;
  if constexpr (W == 8) {
    wym = ld_words<W>(pk, y > 0 ? i - sy : i);
    wzm = ld_words<W>(pk, z > 0 ? i - sz : i);
    wxl = pk[x0 > 0 ? i - 1 : i];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      wym.v[k] = y > 0 ? wym.v[k] : 0u;
      wzm.v[k] = z > 0 ? wzm.v[k] : 0u;
    }
    wxl = x0 > 0 ? wxl : 0u;
  } else {
    if (y > 0) wym = ld_words<W>(pk, i - sy);
    if (z > 0) wzm = ld_words<W>(pk, i - sz);
    if (x0 > 0) wxl = pk[i - 1];
  }
  FV<W> c{}, ym{}, yp{}, zm{}, zp{};
  float xl = 0.0f, xr = 0.0f;
  if (read_p) vec_neighbours<W>(L, x0, y, z, i, ld, c, ym, yp, zm, zp, xl, xr);
  VecStencil<W> st;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const uint32_t u = wo.v[k];
    const float cxp = float(u & 1023u) * L.wx, cyp = float((u >> 10) & 1023u) * L.wy,
                czp = float((u >> 20) & 1023u) * L.wz;
    const float cxm = float((k > 0 ? wo.v[k > 0 ? k - 1 : 0] : wxl) & 1023u) * L.wx;
    const float cym = float((wym.v[k] >> 10) & 1023u) * L.wy, czm = float((wzm.v[k] >> 20) & 1023u) * L.wz;
    const float ck = c.v[k];
    const float left = k > 0 ? c.v[k > 0 ? k - 1 : 0] : xl;
    const float right = k < W - 1 ? c.v[k < W - 1 ? k + 1 : W - 1] : xr;
    float s = 0.0f;
    if (read_p) {
      if (cxp != 0.0f) s += cxp * (right - ck);
      if (cxm != 0.0f) s += cxm * (left - ck);
      if (cyp != 0.0f) s += cyp * (yp.v[k] - ck);
      if (cym != 0.0f) s += cym * (ym.v[k] - ck);
      if (czp != 0.0f) s += czp * (zp.v[k] - ck);
      if (czm != 0.0f) s += czm * (zm.v[k] - ck);
    }
    st.s[k] = s;
    st.d[k] = cxp + cyp + czp + cxm + cym + czm;
    st.c[k] = ck;
    st.f[k] = ((u >> 30) & 1u) ? uint32_t(FLAG_FLUID) : 0u;
  }
  return st;
}

template <int OP, int W, class Ld>
__device__ __forceinline__ VecStencil<W> vec_stencil(const LevelDev& L, int x0, int y, int z, int64_t i, const Ld& ld,
                                                     bool read_p) {
  static_assert(OP != FX_OP_EDGES, "vector stencils cover flag, box and packed levels");
  if constexpr (OP == FX_OP_PACKED) return vec_stencil_packed<W>(L, x0, y, z, i, ld, read_p);
  const int64_t sy = L.nx, sz = int64_t(L.nx) * L.ny;
  VecStencil<W> st;
  vec_flags<W>(L, OP, x0, y, z, i, st.f);
  FV<W> c{}, ym{}, yp{}, zm{}, zp{};
  float xl = 0.0f, xr = 0.0f;
  if (read_p) vec_neighbours<W>(L, x0, y, z, i, ld, c, ym, yp, zm, zp, xl, xr);
  // Fast path, taken when every lane of the wave has all faces of its cells
  // open (the x walls of the lane's two end cells aside): the selects and
  // face counts of flag_cell are known, so only the differences remain.  It
  // evaluates the same operations as flag_cell, bit for bit.
  bool lane_open = true;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const uint32_t need = kAllOpen & ~(k == 0 ? uint32_t(FLAG_XLO) : 0u) & ~(k == W - 1 ? uint32_t(FLAG_XHI) : 0u);
    lane_open = lane_open && (st.f[k] & need) == need;
  }
  if (__all(lane_open)) {
    const float d_open = open_diagonal(L.wx, L.wy, L.wz);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const float ck = c.v[k];
      const float left = k > 0 ? c.v[k > 0 ? k - 1 : 0] : xl;
      const float right = k < W - 1 ? c.v[k < W - 1 ? k + 1 : W - 1] : xr;
      if (k == 0 || k == W - 1) {
        flag_cell(st.f[k], ck, left, right, ym.v[k], yp.v[k], zm.v[k], zp.v[k], L.wx, L.wy, L.wz, st.s[k],
                  st.d[k]);
      } else {
        st.s[k] = open_cell(ck, left, right, ym.v[k], yp.v[k], zm.v[k], zp.v[k], L.wx, L.wy, L.wz);
        st.d[k] = d_open;
      }
      st.c[k] = ck;
    }
    return st;
  }
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const float ck = c.v[k];
    const float left = k > 0 ? c.v[k > 0 ? k - 1 : 0] : xl;
    const float right = k < W - 1 ? c.v[k < W - 1 ? k + 1 : W - 1] : xr;
    flag_cell(st.f[k], ck, left, right, ym.v[k], yp.v[k], zm.v[k], zp.v[k], L.wx, L.wy, L.wz, st.s[k], st.d[k]);
    st.c[k] = ck;
  }
  return st;
}

template <int W>
__device__ __forceinline__ QuadPos vec_pos(const LevelDev& L, int q) {
  const int qx = L.nx / W;
  const int row = fdiv(q, row_div<W>(L));
  QuadPos r;
  r.x = (q - row * qx) * W;
  r.z = fdiv(row, L.ny1);
  r.y = row - r.z * L.ny;
  r.i = int64_t(row) * L.nx + r.x;
  return r;
}

// host: can a level run the quad kernels of the solvers (box or flags
// operator, rows a multiple of the width, 16-byte aligned arrays)?
inline bool quad_ok(const fx_mg_level& L, const void* a = nullptr, const void* b = nullptr,
                    const void* c = nullptr, int width = 4) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  return !op_has_edges(L.op) && L.nx % width == 0 && (L.op != FX_OP_FLAGS || al(L.op_data)) && al(a) && al(b) &&
         al(c);
}

// host: can a level run the vector multigrid kernels (quad_ok, or a packed
// operator with 16-byte aligned words)?
inline bool vec_ok(const fx_mg_level& L, const void* a, const void* b, const void* c, int width) {
  if (L.op != FX_OP_PACKED) return quad_ok(L, a, b, c, width);
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; };
  return L.nx % width == 0 && al(L.op_data) && al(a) && al(b) && al(c);
}

// cells per lane of the level-0 multigrid kernels for a storage type
inline int storage_width(int storage) { return storage == FX_F32 ? 4 : 8; }

}  // namespace fx
