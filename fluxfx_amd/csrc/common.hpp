// Shared device/host helpers for libfluxhip (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/fluxhip.h"

namespace fx {

// ---------------------------------------------------------------------------
// error plumbing: every C-ABI entry returns 0 or a negative code, message kept
// per thread and read back with fx_last_error().
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
int fail_hip(hipError_t e, const char* what);

#define FX_HIP(call)                                                     \
  do {                                                                   \
    hipError_t _e = (call);                                              \
    if (_e != hipSuccess) return ::fx::fail_hip(_e, #call);              \
  } while (0)

#define FX_LAUNCHED(name)                                                \
  do {                                                                   \
    hipError_t _e = hipGetLastError();                                   \
    if (_e != hipSuccess) return ::fx::fail_hip(_e, name);               \
  } while (0)

#define FX_REQUIRE(cond, code, ...)                                      \
  do {                                                                   \
    if (!(cond)) {                                                       \
      ::fx::set_error(__VA_ARGS__);                                      \
      return code;                                                       \
    }                                                                    \
  } while (0)

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

int check_grid(const fx_grid& g);

// ---------------------------------------------------------------------------
// storage types for multigrid levels: fp32, bf16 (raw u16) and fp16
// ---------------------------------------------------------------------------
struct bf16 {
  uint16_t bits;
};

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(bf16 v) { return __uint_as_float(uint32_t(v.bits) << 16); }
__device__ __forceinline__ float to_f(_Float16 v) { return float(v); }

template <typename T>
__device__ __forceinline__ T from_f(float v);
template <>
__device__ __forceinline__ float from_f<float>(float v) { return v; }
template <>
__device__ __forceinline__ bf16 from_f<bf16>(float v) {
  // round to nearest even, NaN stays NaN: gfx950's hardware conversion
  return bf16{__builtin_bit_cast(uint16_t, (__bf16)v)};
}
template <>
__device__ __forceinline__ _Float16 from_f<_Float16>(float v) {
  // the fp32 value is rounded to fp16 on its own: the empty asm keeps the
  // compiler from folding the conversion into the producing fma (a mixed
  // fma rounds once, straight to fp16), so every kernel rounds alike
  asm("" : "+v"(v));
  return _Float16(v);
}

// ---------------------------------------------------------------------------
// indexing helpers (flat x-fast: x + nx*(y + ny*z))
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t idx3(int x, int y, int z, int nx, int ny) {
  return int64_t(x) + int64_t(nx) * (int64_t(y) + int64_t(ny) * int64_t(z));
}

// Per-cell open-face bits of the level-0 operator.
enum : uint8_t {
  FLAG_XLO = 1, FLAG_XHI = 2, FLAG_YLO = 4, FLAG_YHI = 8, FLAG_ZLO = 16, FLAG_ZHI = 32, FLAG_FLUID = 64
};

// Box (unmasked) flags computed from coordinates.
__device__ __forceinline__ uint8_t box_flags(int x, int y, int z, int nx, int ny, int nz) {
  uint8_t f = FLAG_FLUID;
  if (x > 0) f |= FLAG_XLO;
  if (x + 1 < nx) f |= FLAG_XHI;
  if (y > 0) f |= FLAG_YLO;
  if (y + 1 < ny) f |= FLAG_YHI;
  if (z > 0) f |= FLAG_ZLO;
  if (z + 1 < nz) f |= FLAG_ZHI;
  return f;
}

// 7-point operator of one FLAGS/BOX cell in pair form: per axis the two
// neighbour differences t = (p_lo - c) + (p_hi - c) (a closed face adds +0),
// s = wx*tx + wy*ty + wz*tz and d = wx*nx + wy*ny + wz*nz with n the open
// faces per axis.  Every smoother, residual and operator kernel of FLAGS/BOX
// levels (one cell or several per lane) evaluates exactly these operations,
// so their results agree bit for bit; differences are taken before the sums,
// so an offset in p does not cancel catastrophically in fp32.
__device__ __forceinline__ void flag_cell(uint32_t f, float c, float l, float r, float ym, float yp, float zm,
                                          float zp, float wx, float wy, float wz, float& s, float& d) {
  const float tx = ((f & FLAG_XLO) ? l - c : 0.0f) + ((f & FLAG_XHI) ? r - c : 0.0f);
  const float ty = ((f & FLAG_YLO) ? ym - c : 0.0f) + ((f & FLAG_YHI) ? yp - c : 0.0f);
  const float tz = ((f & FLAG_ZLO) ? zm - c : 0.0f) + ((f & FLAG_ZHI) ? zp - c : 0.0f);
  const float nx = float(int((f & FLAG_XLO) != 0) + int((f & FLAG_XHI) != 0));
  const float ny = float(int((f & FLAG_YLO) != 0) + int((f & FLAG_YHI) != 0));
  const float nz = float(int((f & FLAG_ZLO) != 0) + int((f & FLAG_ZHI) != 0));
  s = __builtin_fmaf(wz, tz, __builtin_fmaf(wy, ty, wx * tx));
  d = __builtin_fmaf(wz, nz, __builtin_fmaf(wy, ny, wx * nx));
}

// flag_cell of a cell with all six faces open (the same operations with every
// select taken and n = 2): s only; d is open_diagonal() for every such cell
__device__ __forceinline__ float open_cell(float c, float l, float r, float ym, float yp, float zm, float zp,
                                           float wx, float wy, float wz) {
  const float tx = (l - c) + (r - c), ty = (ym - c) + (yp - c), tz = (zm - c) + (zp - c);
  return __builtin_fmaf(wz, tz, __builtin_fmaf(wy, ty, wx * tx));
}

__device__ __forceinline__ float open_diagonal(float wx, float wy, float wz) {
  return __builtin_fmaf(wz, 2.0f, __builtin_fmaf(wy, 2.0f, wx * 2.0f));
}

// flag_cell of a cell whose four y and z faces are open: s with only the x
// faces tested (the same operations, so the same bits); its diagonal is
// x_diagonal(number of open x faces)
__device__ __forceinline__ float xend_cell(uint32_t f, float c, float l, float r, float ym, float yp, float zm,
                                           float zp, float wx, float wy, float wz) {
  const float tx = ((f & FLAG_XLO) ? l - c : 0.0f) + ((f & FLAG_XHI) ? r - c : 0.0f);
  const float ty = (ym - c) + (yp - c), tz = (zm - c) + (zp - c);
  return __builtin_fmaf(wz, tz, __builtin_fmaf(wy, ty, wx * tx));
}

__device__ __forceinline__ float x_diagonal(int open_x, float wx, float wy, float wz) {
  return __builtin_fmaf(wz, 2.0f, __builtin_fmaf(wy, 2.0f, wx * float(open_x)));
}

constexpr uint32_t kAllOpen = FLAG_XLO | FLAG_XHI | FLAG_YLO | FLAG_YHI | FLAG_ZLO | FLAG_ZHI | FLAG_FLUID;

// FP64 block reduction helpers (wavefront = 64 lanes)
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

// FP64 reduction workspace: ws[q] holds finalised scalar q (q < 16).  Every
// reduction kernel runs on a 1-D grid of exactly kRedBlocks blocks; block b
// writes its partial of quantity q to ws[kWsBase + kRedBlocks*q + b] (no
// atomics, deterministic) and ws_finalize folds the partials into ws[q].
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
constexpr int kWsBase = 16;
constexpr int kWsQuantities = 16;
constexpr int kWsDoubles = kWsBase + kWsQuantities * kRedBlocks;

// Sum `v` over the (1-D, kRedThreads) block and store the block's partial of
// quantity q.  Every thread of the block must call it.
__device__ __forceinline__ void block_partial(double v, double* ws, int q) {
  __shared__ double part[kRedThreads / 64];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < kRedThreads / 64; ++k) s += part[k];
    ws[kWsBase + kRedBlocks * q + blockIdx.x] = s;
  }
  __syncthreads();  // `part` is reused by the next call in the same kernel
}

// clearances above this many cells are stored as 255 (far)
constexpr int kClearanceRadius = 15;

// Active-brick masks of the scalar passes: one byte per 8^3 brick of the cell
// grid, x fastest (bricks at the high ends may be partial).  A plan byte says
// what a pass does with the brick's cells: kBrickSkip leaves them, kBrickCompute
// computes them, kBrickZero stores 0 (outputs known to be exactly zero there).
constexpr int kActiveShift = 3;
constexpr uint8_t kBrickSkip = 0, kBrickCompute = 1, kBrickZero = 2;
struct BrickMask {
  const uint8_t* on;  // null: every brick computed
  int bx, by;         // bricks along x and y
  __device__ __forceinline__ uint32_t at(int x, int y, int z) const {
    return uint32_t(x >> kActiveShift) +
           uint32_t(bx) * (uint32_t(y >> kActiveShift) + uint32_t(by) * uint32_t(z >> kActiveShift));
  }
  __device__ __forceinline__ uint8_t state(int x, int y, int z) const { return on ? on[at(x, y, z)] : kBrickCompute; }
};
inline int active_bricks_along(int n) { return (n + (1 << kActiveShift) - 1) >> kActiveShift; }

// box-restricted mask passes of the moving colliders (advect.hip, pressure.hip);
// b = {x0, y0, z0, x1, y1, z1}, half-open cell ranges inside the grid
int launch_clearance_box(const fx_grid& g, const int* b, const uint8_t* solid, uint8_t* clearance, uint8_t* scratch,
                         uint8_t* scratch2, hipStream_t s);
int launch_flags_box(const fx_grid& g, const int* b, const uint8_t* solid, uint8_t* flags, hipStream_t s);

// host helpers (reduce.hip)
void ws_finalize(double* ws, int q0, int nq, hipStream_t s);
void ws_finalize_mask(double* ws, uint32_t mask, hipStream_t s);
int ws_read(const double* ws, int q0, int nq, double* host, hipStream_t s);

inline dim3 grid3(int nx, int ny, int nz, dim3 block) {
  return dim3((nx + block.x - 1) / block.x, (ny + block.y - 1) / block.y, (nz + block.z - 1) / block.z);
}

}  // namespace fx
