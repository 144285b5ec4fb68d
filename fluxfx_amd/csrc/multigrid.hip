// Geometric multigrid engine: smoothing, fused residual+restriction,
// prolongation, FP64 coarsest direct solve, FP64 norms and the
// tolerance-driven stationary solve.  Templated on the storage type of the
// level arrays (fp32 / bf16 / fp16, arithmetic always fp32) and on the level
// operator (closed box, open-face flags, or edge conductances).
#include "quad.hpp"

namespace fx {

// ---------------------------------------------------------------------------
// smoothing: out = active && d>0 ? p + omega*((s - rhs*inv_dt)/d - p) : 0.
// Every smoother variant evaluates it as the same fused operations, with the
// division as a hardware reciprocal (1 ulp), so they agree bit for bit.
__device__ __forceinline__ float jacobi_update(float c, float s, float b, float d, float inv_dt, float omega) {
  return __builtin_fmaf(omega * __builtin_fmaf(-b, inv_dt, s), __builtin_amdgcn_rcpf(d), c);
}
// zero_guess: p is known to be 0 (fresh coarse level or error equation)
// ---------------------------------------------------------------------------
// One damped-Jacobi update of cell i (correction form of p + omega*((sum w p_n - rhs)/d - p)).
template <typename T, int OP>
__device__ __forceinline__ float smooth_cell(const LevelDev& L, int x, int y, int z, int64_t i, const T* p,
                                             const T* __restrict__ rhs, float inv_dt, float omega, bool read_p) {
  Stencil st = stencil<OP>(L, x, y, z, i, p, read_p);
  float r = 0.0f;
  if (st.active && st.d > 0.0f) r = jacobi_update(st.c, st.s, to_f(rhs[i]), st.d, inv_dt, omega);
  return r;
}

template <typename T, int OP>
__global__ void k_smooth(LevelDev L, const T* __restrict__ p, const T* __restrict__ rhs, T* __restrict__ out,
                         float inv_dt, float omega, int zero_guess) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= L.nx || y >= L.ny) return;
  int64_t i = idx3(x, y, z, L.nx, L.ny);
  out[i] = from_f<T>(smooth_cell<T, OP>(L, x, y, z, i, p, rhs, inv_dt, omega, !zero_guess));
}

// Level-0 smoother for box/flags operators: each thread marches a column of
// `zc` cells along z keeping p[z-1], p[z], p[z+1] in registers; blocks are
// assigned so that each XCD (workgroup id mod 8) sweeps one contiguous slab,
// keeping the y/z neighbour rows in its own L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

template <typename T, int OP>
__global__ __launch_bounds__(256) void k_smooth_zm(LevelDev L, const T* __restrict__ p, const T* __restrict__ rhs,
                                                   T* __restrict__ out, float inv_dt, float omega, int zero_guess,
                                                   int zc, int tiles_x, int tiles_y) {
  int t = xcd_tile(blockIdx.x, gridDim.x);
  int tx = t % tiles_x, rest = t / tiles_x, ty = rest % tiles_y, tz = rest / tiles_y;
  int x = tx * 64 + threadIdx.x, y = ty * 4 + threadIdx.y;
  int z0 = tz * zc, z1 = min(z0 + zc, L.nz);
  if (x >= L.nx || y >= L.ny || z0 >= L.nz) return;
  const int64_t sy = L.nx, sz = int64_t(L.nx) * L.ny;
  int64_t i = idx3(x, y, z0, L.nx, L.ny);
  float pm = 0.0f, pc = 0.0f;
  if (!zero_guess) {
    pc = to_f(p[i]);
    pm = z0 > 0 ? to_f(p[i - sz]) : 0.0f;
  }
  for (int z = z0; z < z1; ++z, i += sz) {
    float pn = (!zero_guess && z + 1 < L.nz) ? to_f(p[i + sz]) : 0.0f;
    uint8_t f = OP == FX_OP_FLAGS ? reinterpret_cast<const uint8_t*>(L.op)[i] : box_flags(x, y, z, L.nx, L.ny, L.nz);
    float n[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (!zero_guess) {
      if (f & FLAG_XLO) n[0] = to_f(p[i - 1]);
      if (f & FLAG_XHI) n[1] = to_f(p[i + 1]);
      if (f & FLAG_YLO) n[2] = to_f(p[i - sy]);
      if (f & FLAG_YHI) n[3] = to_f(p[i + sy]);
    }
    float sdiff, d;
    flag_cell(f, pc, n[0], n[1], n[2], n[3], pm, pn, L.wx, L.wy, L.wz, sdiff, d);
    float r = 0.0f;
    if ((f & FLAG_FLUID) && d > 0.0f) r = jacobi_update(pc, sdiff, to_f(rhs[i]), d, inv_dt, omega);
    out[i] = from_f<T>(r);
    pm = pc;
    pc = pn;
  }
}

// One damped-Jacobi sweep of the W cells of a vector lane.  ZG: p is known to
// be 0 (not read).  On flag/box levels a wave whose lanes have every face of
// their cells open (the x walls of a lane's two end cells aside) takes a path
// without the per-cell selects: every cell is fluid, the inner cells share
// the diagonal of an open cell and the end cells (whose inner x face is open)
// that of 1 or 2 open x faces, so two reciprocals per wave serve all cells.  Both paths evaluate the operations of jacobi_update on the
// values of flag_cell, bit for bit.
template <int OP, int W, bool ZG, class Ld>
__device__ __forceinline__ FV<W> vec_smooth(const LevelDev& L, int x0, int y, int z, int64_t i, const Ld& ld,
                                            const FV<W>& b, float inv_dt, float omega) {
  FV<W> r;
  if constexpr (OP == FX_OP_PACKED) {
    const VecStencil<W> st = vec_stencil_packed<W>(L, x0, y, z, i, ld, !ZG);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const float u = jacobi_update(st.c[k], st.s[k], b.v[k], st.d[k], inv_dt, omega);
      r.v[k] = ((st.f[k] & FLAG_FLUID) && st.d[k] > 0.0f) ? u : 0.0f;
    }
    return r;
  } else {
    uint32_t f[W];
    vec_flags<W>(L, OP, x0, y, z, i, f);
    FV<W> c{}, ym{}, yp{}, zm{}, zp{};
    float xl = 0.0f, xr = 0.0f;
    if (!ZG) vec_neighbours<W>(L, x0, y, z, i, ld, c, ym, yp, zm, zp, xl, xr);
    bool lane_open = true;
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t need = kAllOpen & ~(k == 0 ? uint32_t(FLAG_XLO) : 0u) & ~(k == W - 1 ? uint32_t(FLAG_XHI) : 0u);
      lane_open = lane_open && (f[k] & need) == need;
    }
    const float d_open = open_diagonal(L.wx, L.wy, L.wz);
    if (__all(lane_open) && d_open > 0.0f) {
      const float rd = __builtin_amdgcn_rcpf(d_open);
      const float d1 = x_diagonal(1, L.wx, L.wy, L.wz);
      const float rd1 = __builtin_amdgcn_rcpf(d1);
#pragma unroll
      for (int k = 0; k < W; ++k) {
        const float ck = c.v[k];
        const float left = k > 0 ? c.v[k > 0 ? k - 1 : 0] : xl;
        const float right = k < W - 1 ? c.v[k < W - 1 ? k + 1 : W - 1] : xr;
        if (k == 0 || k == W - 1) {
          const float sk = xend_cell(f[k], ck, left, right, ym.v[k], yp.v[k], zm.v[k], zp.v[k], L.wx, L.wy, L.wz);
          const bool both = (f[k] & (FLAG_XLO | FLAG_XHI)) == (FLAG_XLO | FLAG_XHI);
          const float rk = both ? rd : rd1;
          const bool ok = both || d1 > 0.0f;
          const float u = __builtin_fmaf(omega * __builtin_fmaf(-b.v[k], inv_dt, sk), rk, ck);
          r.v[k] = ok ? u : 0.0f;
        } else {
          const float sk = open_cell(ck, left, right, ym.v[k], yp.v[k], zm.v[k], zp.v[k], L.wx, L.wy, L.wz);
          r.v[k] = __builtin_fmaf(omega * __builtin_fmaf(-b.v[k], inv_dt, sk), rd, ck);
        }
      }
      return r;
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const float ck = c.v[k];
      const float left = k > 0 ? c.v[k > 0 ? k - 1 : 0] : xl;
      const float right = k < W - 1 ? c.v[k < W - 1 ? k + 1 : W - 1] : xr;
      float sk, dk;
      flag_cell(f[k], ck, left, right, ym.v[k], yp.v[k], zm.v[k], zp.v[k], L.wx, L.wy, L.wz, sk, dk);
      const float u = jacobi_update(ck, sk, b.v[k], dk, inv_dt, omega);
      r.v[k] = ((f[k] & FLAG_FLUID) && dk > 0.0f) ? u : 0.0f;
    }
    return r;
  }
}

// Vector smoother (level 0 and other large box/flags/packed levels): one lane
// per W consecutive x cells (one 16-byte load per operand: W = 4 for fp32, 8
// for 16-bit storage), cells enumerated x-fastest over the level; block ids
// are spread so each XCD sweeps a contiguous range of rows.
template <typename T, int OP, int W, bool ZG>
__global__ __launch_bounds__(256) void k_smooth_q(LevelDev L, const T* __restrict__ p, const T* __restrict__ rhs,
                                                  T* __restrict__ out, float inv_dt, float omega, int nq) {
  const int q = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (q >= nq) return;
  const QuadPos at = vec_pos<W>(L, q);
  const FV<W> b = VecIO<T, W>::ld(rhs, at.i);  // issued with the stencil loads
  VecIO<T, W>::st(out, at.i, vec_smooth<OP, W, ZG>(L, at.x, at.y, at.z, at.i, PLoad<T>{p}, b, inv_dt, omega));
}

// Vector residual + 8-child mean: lane u covers fine x W*u .. W*u+W-1 of rows
// 2Y, 2Y+1 and planes 2Z, 2Z+1 and writes the W/2 coarse cells above them
// (bit for bit the sums of k_residual_restrict).
template <typename T, int OP, int W>
__global__ __launch_bounds__(256) void k_residual_restrict_q(LevelDev L, int cny, const T* __restrict__ p,
                                                             const T* __restrict__ rhs, T* __restrict__ crhs,
                                                             float inv_dt, int nu) {
  const int u = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (u >= nu) return;
  const int un = L.nx / W;
  const int crow = fdiv(u, row_div<W>(L)), ux = u - crow * un;
  const int Z = fdiv(crow, L.ny2), Y = crow - Z * cny;
  float acc[W / 2];
#pragma unroll
  for (int j = 0; j < W / 2; ++j) acc[j] = 0.0f;
#pragma unroll 1
  for (int dz = 0; dz < 2; ++dz) {
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
      const int y = 2 * Y + dy, z = 2 * Z + dz, x0 = W * ux;
      const int64_t i = idx3(x0, y, z, L.nx, L.ny);
      const VecStencil<W> st = vec_stencil<OP, W>(L, x0, y, z, i, PLoad<T>{p}, true);
      const FV<W> b = VecIO<T, W>::ld(rhs, i);
      float r[W];
#pragma unroll
      for (int k = 0; k < W; ++k)
        r[k] = ((st.f[k] & FLAG_FLUID) && (!op_has_edges(OP) || st.d[k] > 0.0f)) ? b.v[k] * inv_dt - st.s[k] : 0.0f;
      // one add per child, in the (dz, dy, dx) order of k_residual_restrict
#pragma unroll
      for (int j = 0; j < W / 2; ++j) {
        acc[j] += r[2 * j];
        acc[j] += r[2 * j + 1];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < W / 2; ++j) acc[j] *= 0.125f;
  st_coarse<T, W / 2>(crhs, idx3((W / 2) * ux, Y, Z, (W / 2) * un, cny), acc);
}

// Vector prolongation (in place on p): lane u owns the fine W x 2 x 2 block
// above coarse cells (W/2)u .. (W/2)u + W/2 - 1 of (Y, Z) and reads the clamped
// (W/2 + 2) x 3 x 3 coarse window once.  Per fine cell the eight trilinear
// terms are summed in the same order as k_prolong.  cact: coarse activity bytes
// (renormalised mode on an edge or packed level), else null.  Bit 1 of a
// coarse cell's byte says its whole clamped 3x3x3 neighbourhood is active; when
// it is set for all of a lane's coarse cells every weight sum is exactly 1 and
// the lane skips the activity window (same result, bit for bit).
template <int W, bool RN>
__device__ __forceinline__ void prolong_row(const float (&cv)[3][3][W / 2 + 2], uint64_t amask, int dz, int dy,
                                            float (&corr)[W]) {
  constexpr int CW = W / 2 + 2;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    // even fine x: coarse (x/2 - 1, x/2) with upper weight 3/4; odd: (x/2, x/2 + 1) with 1/4
    const int xb = (k + 1) >> 1;
    const float fz = dz ? 0.25f : 0.75f, fy = dy ? 0.25f : 0.75f, fx = (k & 1) ? 0.25f : 0.75f;
    float num = 0.0f, den = 0.0f;
#pragma unroll
    for (int tz = 0; tz < 2; ++tz) {
      const float wz = tz ? fz : 1.0f - fz;
#pragma unroll
      for (int ty = 0; ty < 2; ++ty) {
        const float wy = ty ? fy : 1.0f - fy;
#pragma unroll
        for (int tx = 0; tx < 2; ++tx) {
          const float wgt = wz * wy * (tx ? fx : 1.0f - fx);
          const float v = cv[dz + tz][dy + ty][xb + tx];
          if (RN) {
            if (!((amask >> (((dz + tz) * 3 + dy + ty) * CW + xb + tx)) & 1u)) continue;
            den += wgt;
          }
          num = __builtin_fmaf(wgt, v, num);
        }
      }
    }
    corr[k] = RN ? (den > 1e-12f ? num / den : 0.0f) : num;
  }
}

// the W/2 consecutive coarse values at an index that is a multiple of W/2
// (8-byte aligned: two fp32 or four 16-bit values) in one load
template <typename T, int N>
__device__ __forceinline__ void ld_span(const T* __restrict__ p, int64_t i, float* out);
template <>
__device__ __forceinline__ void ld_span<float, 2>(const float* __restrict__ p, int64_t i, float* out) {
  const float2 a = *reinterpret_cast<const float2*>(p + i);
  out[0] = a.x;
  out[1] = a.y;
}
template <>
__device__ __forceinline__ void ld_span<bf16, 4>(const bf16* __restrict__ p, int64_t i, float* out) {
  const uint2 a = *reinterpret_cast<const uint2*>(p + i);
  out[0] = __uint_as_float(a.x << 16);
  out[1] = __uint_as_float(a.x & 0xffff0000u);
  out[2] = __uint_as_float(a.y << 16);
  out[3] = __uint_as_float(a.y & 0xffff0000u);
}
template <>
__device__ __forceinline__ void ld_span<_Float16, 4>(const _Float16* __restrict__ p, int64_t i, float* out) {
  typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
  const half4_t a = *reinterpret_cast<const half4_t*>(p + i);
  out[0] = float(a[0]);
  out[1] = float(a[1]);
  out[2] = float(a[2]);
  out[3] = float(a[3]);
}

template <typename T, int FOP, int MODE, int W>
__global__ __launch_bounds__(256) void k_prolong_q(LevelDev F, LevelDev C, const uint8_t* __restrict__ cact,
                                                   T* __restrict__ p, const T* __restrict__ coarse, int nu,
                                                   int zero_in) {
  constexpr int CW = W / 2 + 2;
  static_assert(9 * CW <= 64, "activity window must fit one 64-bit mask");
  const int u = xcd_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (u >= nu) return;
  const int un = F.nx / W;
  const int crow = fdiv(u, row_div<W>(F)), ux = u - crow * un;
  const int Z = fdiv(crow, C.ny1), Y = crow - Z * C.ny;
  const int cx0 = (W / 2) * ux;
  float cv[3][3][CW];
  uint64_t amask = ~uint64_t(0);
  bool renorm = false;
  if (MODE == FX_PROLONG_CONSTANT) {
#pragma unroll
    for (int k = 0; k < W / 2; ++k) cv[1][1][k] = to_f(coarse[idx3(cx0 + k, Y, Z, C.nx, C.ny)]);
  } else {
    int xs[CW];
#pragma unroll
    for (int k = 0; k < CW; ++k) xs[k] = min(max(cx0 - 1 + k, 0), C.nx - 1);
    // away from the x ends the window's middle W/2 values are one aligned load
    const bool inner = cx0 >= 1 && cx0 + W / 2 < C.nx;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int cz = min(max(Z - 1 + a, 0), C.nz - 1);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int cy = min(max(Y - 1 + b, 0), C.ny - 1);
        const int64_t row = (int64_t(cz) * C.ny + cy) * C.nx;
        if (inner) {
          cv[a][b][0] = to_f(coarse[row + cx0 - 1]);
          ld_span<T, W / 2>(coarse, row + cx0, &cv[a][b][1]);
          cv[a][b][CW - 1] = to_f(coarse[row + cx0 + W / 2]);
        } else {
#pragma unroll
          for (int k = 0; k < CW; ++k) cv[a][b][k] = to_f(coarse[row + xs[k]]);
        }
      }
    }
    if (MODE == FX_PROLONG_RENORM && cact) {
      const int64_t own = idx3(cx0, Y, Z, C.nx, C.ny);
      uint32_t all = 2u;
#pragma unroll
      for (int k = 0; k < W / 2; ++k) all &= cact[own + k];
      renorm = all == 0u;
    }
    if (renorm) {
      amask = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int cz = min(max(Z - 1 + a, 0), C.nz - 1);
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const int cy = min(max(Y - 1 + b, 0), C.ny - 1);
          const int64_t row = (int64_t(cz) * C.ny + cy) * C.nx;
#pragma unroll
          for (int k = 0; k < CW; ++k)
            if (cact[row + xs[k]]) amask |= uint64_t(1) << ((a * 3 + b) * CW + k);
        }
      }
    }
  }
#pragma unroll
  for (int dz = 0; dz < 2; ++dz) {
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
      const int y = 2 * Y + dy, z = 2 * Z + dz, x0 = W * ux;
      const int64_t i = idx3(x0, y, z, F.nx, F.ny);
      uint32_t fl[W];
      vec_flags<W>(F, FOP, x0, y, z, i, fl);
      FV<W> r{};
      if (!zero_in) r = VecIO<T, W>::ld(p, i);  // zero_in: p starts from 0 (not read)
      float corr[W];
      if (MODE == FX_PROLONG_CONSTANT) {
#pragma unroll
        for (int k = 0; k < W; ++k) corr[k] = cv[1][1][k >> 1];
      } else if (MODE == FX_PROLONG_RENORM && renorm) {
        prolong_row<W, true>(cv, amask, dz, dy, corr);
      } else {
        prolong_row<W, false>(cv, amask, dz, dy, corr);
      }
#pragma unroll
      for (int k = 0; k < W; ++k) {
        const bool active = (fl[k] & FLAG_FLUID) != 0;
        r.v[k] = (FOP == FX_OP_BOX || active) ? r.v[k] + corr[k] : 0.0f;
      }
      VecIO<T, W>::st(p, i, r);
    }
  }
}

template <typename T, int OP>
__global__ void k_residual(LevelDev L, const T* __restrict__ p, const T* __restrict__ rhs, T* __restrict__ out,
                           float inv_dt) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= L.nx || y >= L.ny) return;
  out[idx3(x, y, z, L.nx, L.ny)] = from_f<T>(residual_at<T, OP>(L, x, y, z, p, rhs, inv_dt));
}

// coarse rhs[C] = mean of the 8 child residuals (shaders/mg_restrict.glsl semantics)
template <typename T, int OP>
__global__ void k_residual_restrict(LevelDev L, int cnx, int cny, const T* __restrict__ p,
                                    const T* __restrict__ rhs, T* __restrict__ crhs, float inv_dt) {
  int X = blockIdx.x * blockDim.x + threadIdx.x, Y = blockIdx.y * blockDim.y + threadIdx.y, Z = blockIdx.z;
  if (X >= cnx || Y >= cny) return;
  float acc = 0.0f;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) acc += residual_at<T, OP>(L, 2 * X + dx, 2 * Y + dy, 2 * Z + dz, p, rhs, inv_dt);
  crhs[idx3(X, Y, Z, cnx, cny)] = from_f<T>(acc * 0.125f);
}

template <typename T>
__global__ void k_restrict_mean(int cnx, int cny, int nx, int ny, const T* __restrict__ fine, T* __restrict__ coarse) {
  int X = blockIdx.x * blockDim.x + threadIdx.x, Y = blockIdx.y * blockDim.y + threadIdx.y, Z = blockIdx.z;
  if (X >= cnx || Y >= cny) return;
  float acc = 0.0f;
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) acc += to_f(fine[idx3(2 * X + dx, 2 * Y + dy, 2 * Z + dz, nx, ny)]);
  coarse[idx3(X, Y, Z, cnx, cny)] = from_f<T>(acc * 0.125f);
}

// Adjoint of clamped trilinear prolongation with unit row sums: per axis
// fine (2i-1, 2i, 2i+1, 2i+2) weighted (1,3,3,1)/8, out-of-range folded onto the edge.
// On masked levels the fine residual is zero in solid cells, and a coarse
// cell with no fluid child (inactive) gets rhs 0.
template <typename T, int COP>
__global__ void k_restrict_full(LevelDev C, int nx, int ny, int nz, const T* __restrict__ fine,
                                T* __restrict__ coarse) {
  const int cnx = C.nx, cny = C.ny;
  int X = blockIdx.x * blockDim.x + threadIdx.x, Y = blockIdx.y * blockDim.y + threadIdx.y, Z = blockIdx.z;
  if (X >= cnx || Y >= cny) return;
  const int64_t ci = idx3(X, Y, Z, cnx, cny);
  if (!cell_active<COP>(C, X, Y, Z, ci)) {
    coarse[ci] = from_f<T>(0.0f);
    return;
  }
  const float wt[4] = {0.125f, 0.375f, 0.375f, 0.125f};
  float acc = 0.0f;
  for (int kz = 0; kz < 4; ++kz) {
    int fz = min(max(2 * Z - 1 + kz, 0), nz - 1);
    for (int ky = 0; ky < 4; ++ky) {
      int fy = min(max(2 * Y - 1 + ky, 0), ny - 1);
      float wzy = wt[kz] * wt[ky];
      for (int kx = 0; kx < 4; ++kx) {
        int fx = min(max(2 * X - 1 + kx, 0), nx - 1);
        acc += wzy * wt[kx] * to_f(fine[idx3(fx, fy, fz, nx, ny)]);
      }
    }
  }
  coarse[ci] = from_f<T>(acc);
}

// prolongation; out may alias in (each fine cell reads only its own value).
// The trilinear terms are explicit fused multiply-adds in a fixed order, so the
// vector kernel (whose products the compiler could otherwise share between
// cells and leave unfused) gives the same bits.
template <typename T, int FOP, int COP, int MODE>
__device__ __forceinline__ float prolong_cell(const LevelDev& F, const LevelDev& C, int x, int y, int z, int64_t i,
                                              float in_val, const T* __restrict__ coarse) {
  if (MODE == FX_PROLONG_CONSTANT) {
    float r = 0.0f;
    if (cell_active<FOP>(F, x, y, z, i)) r = in_val + to_f(coarse[idx3(x >> 1, y >> 1, z >> 1, C.nx, C.ny)]);
    return r;
  }
  // cell-centred trilinear: q = (c+0.5)*0.5-0.5, base = floor(q)
  int bx = (x >> 1) - ((x & 1) ? 0 : 1), by = (y >> 1) - ((y & 1) ? 0 : 1), bz = (z >> 1) - ((z & 1) ? 0 : 1);
  float fx = (x & 1) ? 0.25f : 0.75f, fy = (y & 1) ? 0.25f : 0.75f, fz = (z & 1) ? 0.25f : 0.75f;
  float num = 0.0f, den = 0.0f;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz) {
    int cz = min(max(bz + dz, 0), C.nz - 1);
    float wz = dz ? fz : 1.0f - fz;
#pragma unroll
    for (int dy = 0; dy < 2; ++dy) {
      int cy = min(max(by + dy, 0), C.ny - 1);
      float wy = dy ? fy : 1.0f - fy;
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        int cx = min(max(bx + dx, 0), C.nx - 1);
        float wgt = wz * wy * (dx ? fx : 1.0f - fx);
        int64_t ci = idx3(cx, cy, cz, C.nx, C.ny);
        if (MODE == FX_PROLONG_RENORM) {
          if (!cell_active<COP>(C, cx, cy, cz, ci)) continue;
          den += wgt;
        }
        num = __builtin_fmaf(wgt, to_f(coarse[ci]), num);
      }
    }
  }
  float corr = num;
  if (MODE == FX_PROLONG_RENORM) corr = den > 1e-12f ? num / den : 0.0f;
  float r = in_val + corr;
  if (FOP != FX_OP_BOX && !cell_active<FOP>(F, x, y, z, i)) r = 0.0f;
  return r;
}

template <typename T, int FOP, int COP, int MODE>
__global__ void k_prolong(LevelDev F, LevelDev C, const T* in, const T* __restrict__ coarse, T* out) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= F.nx || y >= F.ny) return;
  int64_t i = idx3(x, y, z, F.nx, F.ny);
  out[i] = from_f<T>(prolong_cell<T, FOP, COP, MODE>(F, C, x, y, z, i, in ? to_f(in[i]) : 0.0f, coarse));
}

// p = scale * M rhs, one wavefront per row, FP64 accumulation
template <typename T>
__global__ void k_direct(int m, const double* __restrict__ M, const T* __restrict__ rhs, T* __restrict__ out,
                         float scale) {
  int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (row >= m) return;
  double acc = 0.0;
  const double* Mr = M + int64_t(row) * m;
  for (int j = lane; j < m; j += 64) acc += Mr[j] * double(to_f(rhs[j]));
  acc = wave_sum(acc);
  if (lane == 0) out[row] = from_f<T>(float(acc * double(scale)));
}

// Coarsest level by conjugate gradients in FP64 inside one workgroup, for
// operators that change between solves (moving colliders), where no
// pseudo-inverse is kept.  One cell per thread, m <= kCgMaxCells.  With A the
// level's operator (A p)_c = sum_n w_n (p_n - p_c), it solves B p = -scale rhs
// for the positive semidefinite B = -A, i.e. A p = scale rhs like the direct
// solve, and returns the pseudo-inverse's solution: B's null space is the
// constants on each connected part of the cells with conductance, so the
// kernel labels the parts (minimum cell index, spread along open faces),
// removes each part's mean from b (the restricted residual's rounding leaves
// some, and CG would blow it up along the null space) and from the result.
// Part sums run over the cells in index order, so results are deterministic.
// CG, preconditioned by the diagonal (cells next to nearly closed coarse faces
// have small diagonals), starts from p = 0 and ends when ||r|| <= 1e-7 ||b||
// (a V-cycle reduces the error by about 4 per cycle, so a coarse correction
// this accurate costs it nothing), when p.Bp <= 0 or
// after `maxit` iterations; every thread computes the same alpha and beta
// from the same LDS partial sums, so control flow stays uniform.  Cells
// without conductance hold 0.  info (nullable): {iterations, ||r|| / ||b||}.
constexpr int kCgMaxCells = 1024;

template <typename T, int OP>
__global__ __launch_bounds__(kCgMaxCells) void k_coarse_cg(LevelDev L, const T* __restrict__ rhs, T* __restrict__ out,
                                                           float scale, int maxit, double* __restrict__ info) {
  __shared__ double P[kCgMaxCells];
  __shared__ double M[kCgMaxCells];
  __shared__ int LB[kCgMaxCells];
  __shared__ double part[2][kCgMaxCells / 64][2];
  const int m = L.nx * L.ny * L.nz;
  const int i = threadIdx.x;
  const int nw = blockDim.x >> 6, lane = i & 63, wv = i >> 6;
  const int sy = L.nx, sz = L.nx * L.ny;
  // conductances to -x, +x, -y, +y, -z, +z (0 where closed or outside)
  double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < m) {
    const int z = i / sz, y = (i - z * sz) / sy, x = i - z * sz - y * sy;
    if (op_has_edges(OP)) {
      // the +x/+y/+z conductances of the level's last cells are wall faces
      // (0); tested anyway, so no neighbour index leaves the level
      const float4 c = edge_at<OP>(L, i);
      w[1] = x + 1 < L.nx ? c.x : 0.0f;
      w[3] = y + 1 < L.ny ? c.y : 0.0f;
      w[5] = z + 1 < L.nz ? c.z : 0.0f;
      w[0] = x > 0 ? edge_x<OP>(L, i - 1) : 0.0f;
      w[2] = y > 0 ? edge_y<OP>(L, i - sy) : 0.0f;
      w[4] = z > 0 ? edge_z<OP>(L, i - sz) : 0.0f;
    } else {
      const uint8_t f =
          OP == FX_OP_FLAGS ? reinterpret_cast<const uint8_t*>(L.op)[i] : box_flags(x, y, z, L.nx, L.ny, L.nz);
      w[0] = (f & FLAG_XLO) ? L.wx : 0.0f;
      w[1] = (f & FLAG_XHI) ? L.wx : 0.0f;
      w[2] = (f & FLAG_YLO) ? L.wy : 0.0f;
      w[3] = (f & FLAG_YHI) ? L.wy : 0.0f;
      w[4] = (f & FLAG_ZLO) ? L.wz : 0.0f;
      w[5] = (f & FLAG_ZHI) ? L.wz : 0.0f;
    }
  }
  // a conductance to a neighbour is that neighbour's conductance back, so the
  // neighbours reached below have conductance themselves
  const int nb[6] = {i - 1, i + 1, i - sy, i + sy, i - sz, i + sz};
  const double d = w[0] + w[1] + w[2] + w[3] + w[4] + w[5];
  const bool live = i < m && d > 0.0;
  // connected parts: every cell takes the smallest label among itself and its
  // open neighbours until no label changes
  int lab = live ? i : -1;
  LB[i] = lab;
  __syncthreads();
  for (;;) {
    int nl = lab;
    if (live) {
#pragma unroll
      for (int k = 0; k < 6; ++k)
        if (w[k] != 0.0) nl = min(nl, LB[nb[k]]);
    }
    const int changed = nl != lab;
    __syncthreads();
    lab = nl;
    LB[i] = lab;
    if (!__syncthreads_or(changed)) break;
  }
  // v minus the mean of v over the cell's part (each part's first cell sums it)
  auto part_free = [&](double v) {
    P[i] = v;
    __syncthreads();
    if (live && lab == i) {
      double s = 0.0;
      int count = 0;
      for (int j = i; j < m; ++j)
        if (LB[j] == i) {
          s += P[j];
          ++count;
        }
      M[i] = s / double(count);
    }
    __syncthreads();
    const double r = live ? v - M[lab] : 0.0;
    __syncthreads();  // P is written again by the next use
    return r;
  };
  // block sums of two values, alternating between the two partial buffers: a
  // buffer is written again only after the barrier of the other buffer's sum,
  // by which time every thread has read it
  int buf = 0;
  auto block_sum2 = [&](double u, double v, double& su, double& sv) {
    u = wave_sum(u);
    v = wave_sum(v);
    if (lane == 0) {
      part[buf][wv][0] = u;
      part[buf][wv][1] = v;
    }
    __syncthreads();
    su = sv = 0.0;
    for (int k = 0; k < nw; ++k) {
      su += part[buf][k][0];
      sv += part[buf][k][1];
    }
    buf ^= 1;
  };
  auto block_sum1 = [&](double u) {
    u = wave_sum(u);
    if (lane == 0) part[buf][wv][0] = u;
    __syncthreads();
    double su = 0.0;
    for (int k = 0; k < nw; ++k) su += part[buf][k][0];
    buf ^= 1;
    return su;
  };
  const double inv_d = live ? 1.0 / d : 0.0;
  const double b = part_free(live ? -double(scale) * double(to_f(rhs[i])) : 0.0);
  double x = 0.0, r = b, z = b * inv_d, p = z;
  double rr, rz;
  block_sum2(r * r, r * z, rr, rz);
  const double bb = rr;
  int it = 0;
  while (it < maxit && rr > 1e-14 * bb) {
    P[i] = p;
    __syncthreads();
    double q = 0.0;
    if (live) {
      q = d * p;
#pragma unroll
      for (int k = 0; k < 6; ++k)
        if (w[k] != 0.0) q -= w[k] * P[nb[k]];
    }
    const double pq = block_sum1(p * q);
    if (!(pq > 0.0)) break;
    const double alpha = rz / pq;
    x += alpha * p;
    r -= alpha * q;
    z = r * inv_d;
    double rz_new;
    block_sum2(r * r, r * z, rr, rz_new);
    ++it;
    const double beta = rz_new / rz;
    rz = rz_new;
    p = z + beta * p;
  }
  __syncthreads();  // the last iteration's P reads are done before part_free writes P
  x = part_free(x);
  if (i < m) out[i] = from_f<T>(live ? float(x) : 0.0f);
  if (info && i == 0) {
    info[0] = double(it);
    info[1] = bb > 0.0 ? sqrt(rr / bb) : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Fused V-cycle legs of the levels below the finest.  A V-cycle visits such a
// level twice: going down it smooths from zero (`pre` sweeps) and restricts
// the residual by the 8-child mean; coming up it adds the prolonged coarse
// correction and smooths `post` times.  Each leg is one launch here instead
// of three: a workgroup owns a kFx x kFy x kFz block of fine cells, stages its
// operands with a halo into LDS and recomputes the halo's sweeps itself
// (temporal blocking).  Values are rounded to the storage type exactly where
// the unfused path stores them and each cell's arithmetic is that of the
// one-cell kernels, so both paths give the same bits.  Cells outside the
// level hold 0 in every staged array (zero conductances, inactive), which the
// one-cell kernels' bounds tests amount to.
constexpr int kFx = 8, kFy = 8, kFz = 8;  // fine cells per workgroup
constexpr int kLegThreads = 512;

// k over [0, N) in kLegThreads strides, unrolled so that a thread's loads of
// a staging pass are all in flight together
#define FX_TILE_FOR(N, k)                                                                 \
  _Pragma("unroll") for (int j_ = 0; j_ < ((N) + kLegThreads - 1) / kLegThreads; ++j_)   \
    if (const int k = int(threadIdx.x) + j_ * kLegThreads; k < (N))

// A box of tile-relative cells [-H, k + H) per axis, x fastest.
template <int H>
struct TileBox {
  static constexpr int ex = kFx + 2 * H, ey = kFy + 2 * H, ez = kFz + 2 * H, n = ex * ey * ez;
  static __device__ __forceinline__ int at(int x, int y, int z) { return (x + H) + ex * ((y + H) + ey * (z + H)); }
  static __device__ __forceinline__ void coords(int k, int& x, int& y, int& z) {
    x = k % ex - H;
    y = (k / ex) % ey - H;
    z = k / (ex * ey) - H;
  }
};

// 7-point operator of a BOX or PACKED cell from staged words (halo HW) and
// staged p (halo HP): the operations of stencil<OP>.
template <int OP, int HW, int HP>
__device__ __forceinline__ Stencil tile_stencil(const LevelDev& L, const uint32_t* wd, const float* pv, int x, int y,
                                                int z, int gx, int gy, int gz, bool read_p) {
  using BW = TileBox<HW>;
  using BP = TileBox<HP>;
  Stencil st{0.0f, 0.0f, 0.0f, true};
  const float c0 = read_p ? pv[BP::at(x, y, z)] : 0.0f;
  st.c = c0;
  if constexpr (OP == FX_OP_PACKED) {
    const uint32_t u = wd[BW::at(x, y, z)];
    const float cx = float(u & 1023u) * L.wx, cy = float((u >> 10) & 1023u) * L.wy,
                cz = float((u >> 20) & 1023u) * L.wz;
    const float wxl = float(wd[BW::at(x - 1, y, z)] & 1023u) * L.wx;
    const float wyl = float((wd[BW::at(x, y - 1, z)] >> 10) & 1023u) * L.wy;
    const float wzl = float((wd[BW::at(x, y, z - 1)] >> 20) & 1023u) * L.wz;
    st.d = cx + cy + cz + wxl + wyl + wzl;
    st.active = (u >> 30) & 1u;
    if (read_p) {
      float s = 0.0f;
      if (cx != 0.0f) s += cx * (pv[BP::at(x + 1, y, z)] - c0);
      if (wxl != 0.0f) s += wxl * (pv[BP::at(x - 1, y, z)] - c0);
      if (cy != 0.0f) s += cy * (pv[BP::at(x, y + 1, z)] - c0);
      if (wyl != 0.0f) s += wyl * (pv[BP::at(x, y - 1, z)] - c0);
      if (cz != 0.0f) s += cz * (pv[BP::at(x, y, z + 1)] - c0);
      if (wzl != 0.0f) s += wzl * (pv[BP::at(x, y, z - 1)] - c0);
      st.s = s;
    }
  } else {
    static_assert(OP == FX_OP_BOX, "fused legs cover box and packed levels");
    const uint8_t f = box_flags(gx, gy, gz, L.nx, L.ny, L.nz);
    float n[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (read_p) {
      if (f & FLAG_XLO) n[0] = pv[BP::at(x - 1, y, z)];
      if (f & FLAG_XHI) n[1] = pv[BP::at(x + 1, y, z)];
      if (f & FLAG_YLO) n[2] = pv[BP::at(x, y - 1, z)];
      if (f & FLAG_YHI) n[3] = pv[BP::at(x, y + 1, z)];
      if (f & FLAG_ZLO) n[4] = pv[BP::at(x, y, z - 1)];
      if (f & FLAG_ZHI) n[5] = pv[BP::at(x, y, z + 1)];
    }
    flag_cell(f, c0, n[0], n[1], n[2], n[3], n[4], n[5], L.wx, L.wy, L.wz, st.s, st.d);
  }
  return st;
}

__device__ __forceinline__ bool in_level(const LevelDev& L, int gx, int gy, int gz) {
  return gx >= 0 && gy >= 0 && gz >= 0 && gx < L.nx && gy < L.ny && gz < L.nz;
}

// A level array on TileBox<H> staged through registers into LDS laid out on
// TileBox<HL> (HL >= H), 0 outside the level.  load() only issues loads
// (out-of-level lanes read element 0 and discard it), store() writes LDS: a
// kernel issues the loads of all its arrays before the first store, so one
// memory round trip covers its staging.
template <int H, int HL, typename T>
struct Staged {
  using B = TileBox<H>;
  static constexpr int NJ = (B::n + kLegThreads - 1) / kLegThreads;
  T v[NJ];
  bool ok[NJ];
  int at[NJ];
  __device__ __forceinline__ void load(const LevelDev& L, int x0, int y0, int z0, const T* __restrict__ src) {
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int k = int(threadIdx.x) + j * kLegThreads;
      int x, y, z;
      B::coords(k < B::n ? k : B::n - 1, x, y, z);
      const int gx = x0 + x, gy = y0 + y, gz = z0 + z;
      ok[j] = k < B::n && in_level(L, gx, gy, gz);
      at[j] = k < B::n ? TileBox<HL>::at(x, y, z) : -1;
      v[j] = src[ok[j] ? idx3(gx, gy, gz, L.nx, L.ny) : 0];
    }
  }
  __device__ __forceinline__ void store_float(float* dst) const {
#pragma unroll
    for (int j = 0; j < NJ; ++j)
      if (at[j] >= 0) dst[at[j]] = ok[j] ? to_f(v[j]) : 0.0f;
  }
  __device__ __forceinline__ void store_word(uint32_t* dst) const {
#pragma unroll
    for (int j = 0; j < NJ; ++j)
      if (at[j] >= 0) dst[at[j]] = ok[j] ? uint32_t(v[j]) : 0u;
  }
};

// One damped-Jacobi sweep over the cells of TileBox<HO> from p (read_p) or
// from zero, rounded to T; p, the result and rhs (rb) are laid out on
// TileBox<HL>, HL >= HO + 1, the words on TileBox<HW>.
template <typename T, int OP, int HW, int HL, int HO>
__device__ __forceinline__ void tile_sweep(const LevelDev& L, int x0, int y0, int z0, const uint32_t* wd,
                                           const float* pin, const float* rb, float* pout, bool read_p, float inv_dt,
                                           float omega) {
  using BO = TileBox<HO>;
  using BL = TileBox<HL>;
  FX_TILE_FOR(BO::n, k) {
    int x, y, z;
    BO::coords(k, x, y, z);
    const int gx = x0 + x, gy = y0 + y, gz = z0 + z;
    float r = 0.0f;
    if (in_level(L, gx, gy, gz)) {
      const Stencil st = tile_stencil<OP, HW, HL>(L, wd, pin, x, y, z, gx, gy, gz, read_p);
      if (st.active && st.d > 0.0f)
        r = to_f(from_f<T>(jacobi_update(st.c, st.s, rb[BL::at(x, y, z)], st.d, inv_dt, omega)));
    }
    pout[BL::at(x, y, z)] = r;
  }
}

// Down leg: S sweeps from zero, written to p; the 8-child mean of the
// residual after them, written to the coarse rhs.  Sweep j runs on the tile
// with a halo of S - j + 1 cells; p arrays and rhs are laid out on halo S.
template <typename T, int OP, int S>
__global__ __launch_bounds__(kLegThreads) void k_leg_down(LevelDev L, int tiles_x, int tiles_y,
                                                          const T* __restrict__ rhs, T* __restrict__ p,
                                                          T* __restrict__ crhs, float inv_dt, float omega) {
  static_assert(S >= 1 && S <= 3, "the fused legs stage one to three sweeps");
  using BW = TileBox<S + 1>;
  using BA = TileBox<S>;
  using BR = TileBox<0>;
  __shared__ uint32_t wd[OP == FX_OP_PACKED ? BW::n : 1];
  __shared__ float rb[BA::n];
  __shared__ float pa[BA::n];
  __shared__ float pb[S >= 2 ? BA::n : 1];
  __shared__ float res[BR::n];
  const int t = blockIdx.x;
  const int x0 = (t % tiles_x) * kFx, y0 = ((t / tiles_x) % tiles_y) * kFy, z0 = (t / tiles_x / tiles_y) * kFz;
  Staged<S + 1, S + 1, uint32_t> sw;
  Staged<S, S, T> sr;
  if constexpr (OP == FX_OP_PACKED) sw.load(L, x0, y0, z0, reinterpret_cast<const uint32_t*>(L.op));
  sr.load(L, x0, y0, z0, rhs);
  if constexpr (OP == FX_OP_PACKED) sw.store_word(wd);
  sr.store_float(rb);
  __syncthreads();
  tile_sweep<T, OP, S + 1, S, S>(L, x0, y0, z0, wd, pa, rb, pa, false, inv_dt, omega);
  __syncthreads();
  if constexpr (S >= 2) {
    tile_sweep<T, OP, S + 1, S, S - 1>(L, x0, y0, z0, wd, pa, rb, pb, true, inv_dt, omega);
    __syncthreads();
  }
  if constexpr (S >= 3) {
    tile_sweep<T, OP, S + 1, S, S - 2>(L, x0, y0, z0, wd, pb, rb, pa, true, inv_dt, omega);
    __syncthreads();
  }
  const float* pf = (S % 2 == 1) ? pa : pb;  // the last sweep's output (halo 1)
  FX_TILE_FOR(BR::n, k) {
    int x, y, z;
    BR::coords(k, x, y, z);
    const int gx = x0 + x, gy = y0 + y, gz = z0 + z;
    float r = 0.0f;
    if (in_level(L, gx, gy, gz)) {
      p[idx3(gx, gy, gz, L.nx, L.ny)] = from_f<T>(pf[BA::at(x, y, z)]);
      // residual_at: rhs*inv_dt - s on active cells
      const Stencil st = tile_stencil<OP, S + 1, S>(L, wd, pf, x, y, z, gx, gy, gz, true);
      if (st.active && (OP != FX_OP_PACKED || st.d > 0.0f)) r = rb[BA::at(x, y, z)] * inv_dt - st.s;
    }
    res[k] = r;
  }
  __syncthreads();
  const int cnx = L.nx / 2, cny = L.ny / 2, cnz = L.nz / 2;
  FX_TILE_FOR((kFx / 2) * (kFy / 2) * (kFz / 2), k) {
    const int cx = k % (kFx / 2), cy = (k / (kFx / 2)) % (kFy / 2), cz = k / ((kFx / 2) * (kFy / 2));
    const int X = x0 / 2 + cx, Y = y0 / 2 + cy, Z = z0 / 2 + cz;
    if (X >= cnx || Y >= cny || Z >= cnz) continue;
    float acc = 0.0f;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) acc += res[BR::at(2 * cx + dx, 2 * cy + dy, 2 * cz + dz)];
    crhs[idx3(X, Y, Z, cnx, cny)] = from_f<T>(acc * 0.125f);
  }
}

// Up leg: p + prolonged coarse correction (prolong_cell's operations) on the
// tile with a halo of S cells, then S sweeps (sweep j on halo S - j), the
// last one written to out (out and p are distinct arrays).
template <typename T, int OP, int MODE, int S>
__global__ __launch_bounds__(kLegThreads) void k_leg_up(LevelDev F, LevelDev C, const uint8_t* __restrict__ cact, int tiles_x, int tiles_y,
                                                        const T* __restrict__ p, const T* __restrict__ coarse,
                                                        const T* __restrict__ rhs, T* __restrict__ out, float inv_dt,
                                                        float omega) {
  static_assert(S >= 1 && S <= 3, "the fused legs stage one to three sweeps");
  using BU = TileBox<S>;
  using BV = TileBox<S - 1>;
  using BR = TileBox<0>;
  // coarse window [x0/2 - 2, x0/2 + kFx/2 + 2) per axis: the clamped trilinear
  // neighbours of every fine cell within 3 cells of the tile
  constexpr int CWX = kFx / 2 + 4, CWY = kFy / 2 + 4, CWZ = kFz / 2 + 4, CWN = CWX * CWY * CWZ;
  __shared__ uint32_t wd[OP == FX_OP_PACKED ? BU::n : 1];
  __shared__ float cw[CWN];
  __shared__ uint8_t ca[MODE == FX_PROLONG_RENORM ? CWN : 1];
  __shared__ float rb[BU::n];
  __shared__ float pa[BU::n];
  __shared__ float pb[S >= 2 ? BU::n : 1];
  const int t = blockIdx.x;
  const int x0 = (t % tiles_x) * kFx, y0 = ((t / tiles_x) % tiles_y) * kFy, z0 = (t / tiles_x / tiles_y) * kFz;
  const int wx0 = x0 / 2 - 2, wy0 = y0 / 2 - 2, wz0 = z0 / 2 - 2;
  Staged<S, S, uint32_t> sw;
  Staged<S - 1, S, T> sr;
  Staged<S, S, T> sp;
  if constexpr (OP == FX_OP_PACKED) sw.load(F, x0, y0, z0, reinterpret_cast<const uint32_t*>(F.op));
  sr.load(F, x0, y0, z0, rhs);
  sp.load(F, x0, y0, z0, p);
  // the coarse window with clamped indices (every index valid)
  float cv[(CWN + kLegThreads - 1) / kLegThreads];
  uint32_t cflag[(CWN + kLegThreads - 1) / kLegThreads];
#pragma unroll
  for (int j = 0; j < (CWN + kLegThreads - 1) / kLegThreads; ++j) {
    const int k = min(int(threadIdx.x) + j * kLegThreads, CWN - 1);
    const int cx = min(max(wx0 + k % CWX, 0), C.nx - 1), cy = min(max(wy0 + (k / CWX) % CWY, 0), C.ny - 1),
              cz = min(max(wz0 + k / (CWX * CWY), 0), C.nz - 1);
    const int64_t ci = idx3(cx, cy, cz, C.nx, C.ny);
    cv[j] = to_f(coarse[ci]);
    cflag[j] = (MODE == FX_PROLONG_RENORM && cact) ? cact[ci] : 1u;
  }
  if constexpr (OP == FX_OP_PACKED) sw.store_word(wd);
  sr.store_float(rb);
  sp.store_float(pa);
#pragma unroll
  for (int j = 0; j < (CWN + kLegThreads - 1) / kLegThreads; ++j) {
    const int k = int(threadIdx.x) + j * kLegThreads;
    if (k < CWN) {
      cw[k] = cv[j];
      if (MODE == FX_PROLONG_RENORM) ca[k] = cflag[j] & 1u;
    }
  }
  __syncthreads();
  FX_TILE_FOR(BU::n, k) {
    int x, y, z;
    BU::coords(k, x, y, z);
    const int gx = x0 + x, gy = y0 + y, gz = z0 + z;
    float r = 0.0f;
    if (in_level(F, gx, gy, gz)) {
      const int bx = (gx >> 1) - ((gx & 1) ? 0 : 1), by = (gy >> 1) - ((gy & 1) ? 0 : 1),
                bz = (gz >> 1) - ((gz & 1) ? 0 : 1);
      const float fx = (gx & 1) ? 0.25f : 0.75f, fy = (gy & 1) ? 0.25f : 0.75f, fz = (gz & 1) ? 0.25f : 0.75f;
      float num = 0.0f, den = 0.0f;
#pragma unroll
      for (int dz = 0; dz < 2; ++dz) {
        const float wz = dz ? fz : 1.0f - fz;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
          const float wy = dy ? fy : 1.0f - fy;
#pragma unroll
          for (int dx = 0; dx < 2; ++dx) {
            const float wgt = wz * wy * (dx ? fx : 1.0f - fx);
            const int j = (bx + dx - wx0) + CWX * ((by + dy - wy0) + CWY * (bz + dz - wz0));
            if (MODE == FX_PROLONG_RENORM) {
              if (!ca[j]) continue;
              den += wgt;
            }
            num = __builtin_fmaf(wgt, cw[j], num);
          }
        }
      }
      float corr = num;
      if (MODE == FX_PROLONG_RENORM) corr = den > 1e-12f ? num / den : 0.0f;
      r = pa[k] + corr;
      if (OP == FX_OP_PACKED && !((wd[BU::at(x, y, z)] >> 30) & 1u)) r = 0.0f;
      r = to_f(from_f<T>(r));
    }
    pa[k] = r;
  }
  __syncthreads();
  if constexpr (S >= 2) {
    tile_sweep<T, OP, S, S, S - 1>(F, x0, y0, z0, wd, pa, rb, pb, true, inv_dt, omega);
    __syncthreads();
  }
  if constexpr (S >= 3) {
    tile_sweep<T, OP, S, S, S - 2>(F, x0, y0, z0, wd, pb, rb, pa, true, inv_dt, omega);
    __syncthreads();
  }
  const float* pf = (S % 2 == 1) ? pa : pb;  // input of the last sweep (halo 1)
  FX_TILE_FOR(BR::n, k) {
    int x, y, z;
    BR::coords(k, x, y, z);
    const int gx = x0 + x, gy = y0 + y, gz = z0 + z;
    if (!in_level(F, gx, gy, gz)) continue;
    const Stencil st = tile_stencil<OP, S, S>(F, wd, pf, x, y, z, gx, gy, gz, true);
    float r = 0.0f;
    if (st.active && st.d > 0.0f) r = jacobi_update(st.c, st.s, rb[BU::at(x, y, z)], st.d, inv_dt, omega);
    out[idx3(gx, gy, gz, F.nx, F.ny)] = from_f<T>(r);
  }
}

static const dim3 kBlock = kMgBlock;

// ---------------------------------------------------------------------------
// host-side dispatch over (storage, operator)
// ---------------------------------------------------------------------------
int check_config(const fx_mg_config* c) {
  FX_REQUIRE(c, FX_EINVAL, "null config");
  FX_REQUIRE(c->levels >= 1 && c->levels <= 16, FX_EINVAL, "levels must be in [1,16]");
  FX_REQUIRE(c->storage >= 0 && c->storage <= 2, FX_EINVAL, "bad storage type");
  for (int l = 0; l < c->levels; ++l) {
    const fx_mg_level& L = c->level[l];
    FX_REQUIRE(L.nx >= 1 && L.ny >= 1 && L.nz >= 1, FX_EINVAL, "level %d has an empty shape", l);
    FX_REQUIRE(L.p && L.q && L.rhs && L.p != L.q, FX_EINVAL, "level %d needs distinct p, q and rhs", l);
    FX_REQUIRE(L.op == FX_OP_BOX || L.op_data, FX_EINVAL, "level %d operator data missing", l);
    if (l > 0) {
      const fx_mg_level& F = c->level[l - 1];
      FX_REQUIRE(F.nx == 2 * L.nx && F.ny == 2 * L.ny && F.nz == 2 * L.nz, FX_EINVAL,
                 "level %d is not an exact factor-two coarsening", l);
    }
  }
  const fx_mg_level& C = c->level[c->levels - 1];
  if (c->coarse_sweeps == 0) {
    FX_REQUIRE(c->coarse_matrix, FX_EINVAL, "direct coarse solve needs a matrix");
  } else if (c->coarse_sweeps < 0) {
    FX_REQUIRE(C.nx * C.ny * C.nz <= kCgMaxCells, FX_EINVAL,
               "coarse conjugate gradients limited to %d cells, got %d", kCgMaxCells, C.nx * C.ny * C.nz);
  }
  (void)C;
  return FX_OK;
}

static int launch_smooth(fx_mg_config* c, int l, float inv_dt, bool zero_guess, hipStream_t s) {
  fx_mg_level& L = c->level[l];
  LevelDev D = dev_level(L);
  const int64_t cells = int64_t(L.nx) * L.ny * L.nz;
  const int W = storage_width(c->storage);
  if (cells >= kQuadMinCells && vec_ok(L, L.p, L.q, L.rhs, W)) {
    const int nq = int(cells / W);
    dim3 grid((nq + 255) / 256);
#define FX_SMQ(OPC)                                                                                          \
  do {                                                                                                        \
    if (zero_guess)                                                                                           \
      hipLaunchKernelGGL((k_smooth_q<TV, OPC, WV, true>), grid, dim3(256), 0, s, D, (const TV*)L.p,           \
                         (const TV*)L.rhs, (TV*)L.q, inv_dt, c->omega, nq);                                   \
    else                                                                                                      \
      hipLaunchKernelGGL((k_smooth_q<TV, OPC, WV, false>), grid, dim3(256), 0, s, D, (const TV*)L.p,          \
                         (const TV*)L.rhs, (TV*)L.q, inv_dt, c->omega, nq);                                   \
  } while (0)
    FX_DISPATCH_T(c->storage, {
      constexpr int WV = vec_width<TV>();
      if (L.op == FX_OP_FLAGS) FX_SMQ(FX_OP_FLAGS);
      else if (L.op == FX_OP_PACKED) FX_SMQ(FX_OP_PACKED);
      else FX_SMQ(FX_OP_BOX);
    });
#undef FX_SMQ
  } else if (!op_has_edges(L.op) && cells >= (1 << 18)) {
    const int zc = 16;
    int tx = (L.nx + 63) / 64, ty = (L.ny + 3) / 4, tz = (L.nz + zc - 1) / zc;
    dim3 grid(tx * ty * tz);
    FX_DISPATCH_T(c->storage, {
      if (L.op == FX_OP_FLAGS)
        hipLaunchKernelGGL((k_smooth_zm<TV, FX_OP_FLAGS>), grid, kBlock, 0, s, D, (const TV*)L.p, (const TV*)L.rhs,
                           (TV*)L.q, inv_dt, c->omega, zero_guess ? 1 : 0, zc, tx, ty);
      else
        hipLaunchKernelGGL((k_smooth_zm<TV, FX_OP_BOX>), grid, kBlock, 0, s, D, (const TV*)L.p, (const TV*)L.rhs,
                           (TV*)L.q, inv_dt, c->omega, zero_guess ? 1 : 0, zc, tx, ty);
    });
  } else {
    FX_DISPATCH_T(c->storage, FX_DISPATCH_OP(L.op, {
      hipLaunchKernelGGL((k_smooth<TV, OPV>), level_grid(L), kBlock, 0, s, D, (const TV*)L.p, (const TV*)L.rhs,
                         (TV*)L.q, inv_dt, c->omega, zero_guess ? 1 : 0);
    }));
  }
  FX_LAUNCHED("k_smooth");
  void* t = L.p;
  L.p = L.q;
  L.q = t;
  return FX_OK;
}

static int launch_fill_zero(fx_mg_config* c, void* ptr, int64_t n, hipStream_t s);

static int smooth_n(fx_mg_config* c, int l, int count, float inv_dt, bool zero_guess, hipStream_t s) {
  for (int k = 0; k < count; ++k) {
    if (int e = launch_smooth(c, l, inv_dt, zero_guess && k == 0, s)) return e;
  }
  return FX_OK;
}

static int launch_residual_restrict(fx_mg_config* c, int l, float inv_dt, hipStream_t s) {
  fx_mg_level& F = c->level[l];
  fx_mg_level& C = c->level[l + 1];
  LevelDev D = dev_level(F);
  const int64_t fcells = int64_t(F.nx) * F.ny * F.nz;
  const int W = storage_width(c->storage);
  if (c->restriction == FX_RESTRICT_MEAN && fcells >= kQuadMinCells && vec_ok(F, F.p, F.rhs, C.rhs, W)) {
    const int nu = int(int64_t(F.nx / W) * C.ny * C.nz);
    dim3 grid((nu + 255) / 256);
#define FX_RRQ(OPC)                                                                                         \
  hipLaunchKernelGGL((k_residual_restrict_q<TV, OPC, WV>), grid, dim3(256), 0, s, D, C.ny, (const TV*)F.p, \
                     (const TV*)F.rhs, (TV*)C.rhs, inv_dt, nu)
    FX_DISPATCH_T(c->storage, {
      constexpr int WV = vec_width<TV>();
      if (F.op == FX_OP_FLAGS) FX_RRQ(FX_OP_FLAGS);
      else if (F.op == FX_OP_PACKED) FX_RRQ(FX_OP_PACKED);
      else FX_RRQ(FX_OP_BOX);
    });
#undef FX_RRQ
    FX_LAUNCHED("k_residual_restrict_q");
  } else if (c->restriction == FX_RESTRICT_MEAN) {
    FX_DISPATCH_T(c->storage, FX_DISPATCH_OP(F.op, {
      hipLaunchKernelGGL((k_residual_restrict<TV, OPV>), level_grid(C), kBlock, 0, s, D, C.nx, C.ny,
                         (const TV*)F.p, (const TV*)F.rhs, (TV*)C.rhs, inv_dt);
    }));
    FX_LAUNCHED("k_residual_restrict");
  } else {
    // residual into the free ping-pong partner, then the 64-point restriction
    FX_DISPATCH_T(c->storage, FX_DISPATCH_OP(F.op, {
      hipLaunchKernelGGL((k_residual<TV, OPV>), level_grid(F), kBlock, 0, s, D, (const TV*)F.p,
                         (const TV*)F.rhs, (TV*)F.q, inv_dt);
    }));
    FX_LAUNCHED("k_residual");
    const LevelDev DC = dev_level(C);
    FX_DISPATCH_T(c->storage, FX_DISPATCH_OP(C.op, {
      hipLaunchKernelGGL((k_restrict_full<TV, OPV>), level_grid(C), kBlock, 0, s, DC, F.nx, F.ny, F.nz,
                         (const TV*)F.q, (TV*)C.rhs);
    }));
    FX_LAUNCHED("k_restrict_full");
  }
  return FX_OK;
}

static int launch_prolong(const fx_mg_config* c, int l, const void* in, const void* coarse, void* out,
                          hipStream_t s) {
  const fx_mg_level& F = c->level[l];
  const fx_mg_level& C = c->level[l + 1];
  LevelDev DF = dev_level(F), DC = dev_level(C);
  int mode = c->prolongation;
  FX_REQUIRE(mode >= 0 && mode <= 2, FX_EINVAL, "bad prolongation mode");
  FX_REQUIRE(F.op != FX_OP_FLAGS || C.op != FX_OP_FLAGS, FX_EINVAL, "flags operators are level-0 only");
  const int64_t fcells = int64_t(F.nx) * F.ny * F.nz;
  const bool coarse_known = mode != FX_PROLONG_RENORM || C.op == FX_OP_BOX || (op_has_edges(C.op) && C.active);
  const int W = storage_width(c->storage);
  // the vector kernel reads W/2 coarse values at a time: an 8-byte aligned coarse array
  const bool coarse_aligned = (reinterpret_cast<uintptr_t>(coarse) & 7u) == 0;
  if ((in == out || !in) && fcells >= kQuadMinCells && coarse_known && C.op != FX_OP_FLAGS && coarse_aligned &&
      vec_ok(F, out, nullptr, nullptr, W)) {
    const int zero_in = in ? 0 : 1;
    const int nu = int(int64_t(F.nx / W) * C.ny * C.nz);
    dim3 grid((nu + 255) / 256);
    const uint8_t* cact = op_has_edges(C.op) ? C.active : nullptr;
#define FX_PROQ(FO, MO)                                                                                 \
  FX_DISPATCH_T(c->storage, {                                                                           \
    constexpr int WV = vec_width<TV>();                                                                 \
    hipLaunchKernelGGL((k_prolong_q<TV, FO, MO, WV>), grid, dim3(256), 0, s, DF, DC, cact, (TV*)out,    \
                       (const TV*)coarse, nu, zero_in);                                                 \
  })
    if (F.op == FX_OP_FLAGS) {
      if (mode == FX_PROLONG_CONSTANT) { FX_PROQ(FX_OP_FLAGS, FX_PROLONG_CONSTANT); }
      else if (mode == FX_PROLONG_TRILINEAR) { FX_PROQ(FX_OP_FLAGS, FX_PROLONG_TRILINEAR); }
      else { FX_PROQ(FX_OP_FLAGS, FX_PROLONG_RENORM); }
    } else if (F.op == FX_OP_PACKED) {
      if (mode == FX_PROLONG_CONSTANT) { FX_PROQ(FX_OP_PACKED, FX_PROLONG_CONSTANT); }
      else if (mode == FX_PROLONG_TRILINEAR) { FX_PROQ(FX_OP_PACKED, FX_PROLONG_TRILINEAR); }
      else { FX_PROQ(FX_OP_PACKED, FX_PROLONG_RENORM); }
    } else {
      if (mode == FX_PROLONG_CONSTANT) { FX_PROQ(FX_OP_BOX, FX_PROLONG_CONSTANT); }
      else if (mode == FX_PROLONG_TRILINEAR) { FX_PROQ(FX_OP_BOX, FX_PROLONG_TRILINEAR); }
      else { FX_PROQ(FX_OP_BOX, FX_PROLONG_RENORM); }
    }
#undef FX_PROQ
    FX_LAUNCHED("k_prolong_q");
    return FX_OK;
  }
#define FX_PRO(FO, CO, MO)                                                                             \
  FX_DISPATCH_T(c->storage, {                                                                          \
    hipLaunchKernelGGL((k_prolong<TV, FO, CO, MO>), level_grid(F), kBlock, 0, s, DF, DC, (const TV*)in, \
                       (const TV*)coarse, (TV*)out);                                                   \
  })
  // the coarse operator matters only for the activity test of renormalised
  // prolongation; piecewise-constant prolongation tests the fine cell only
  if (mode == FX_PROLONG_CONSTANT) {
    FX_DISPATCH_OP(F.op, { FX_PRO(OPV, FX_OP_BOX, FX_PROLONG_CONSTANT); });
  } else if (mode == FX_PROLONG_TRILINEAR) {
    FX_DISPATCH_OP(F.op, { FX_PRO(OPV, FX_OP_BOX, FX_PROLONG_TRILINEAR); });
  } else {
    FX_REQUIRE(C.op != FX_OP_FLAGS, FX_EINVAL, "renormalised prolongation needs a coarse edge or box level");
    if (C.op == FX_OP_EDGES) {
      FX_DISPATCH_OP(F.op, { FX_PRO(OPV, FX_OP_EDGES, FX_PROLONG_RENORM); });
    } else if (C.op == FX_OP_PACKED) {
      FX_DISPATCH_OP(F.op, { FX_PRO(OPV, FX_OP_PACKED, FX_PROLONG_RENORM); });
    } else {
      FX_DISPATCH_OP(F.op, { FX_PRO(OPV, FX_OP_BOX, FX_PROLONG_RENORM); });
    }
  }
#undef FX_PRO
  FX_LAUNCHED("k_prolong");
  return FX_OK;
}

static int launch_coarse_cg(const fx_mg_config* c, double* info, hipStream_t s) {
  const fx_mg_level& C = c->level[c->levels - 1];
  const int m = C.nx * C.ny * C.nz;
  FX_REQUIRE(m <= kCgMaxCells, FX_EINVAL, "coarse conjugate gradients limited to %d cells, got %d", kCgMaxCells, m);
  const dim3 block((m + 63) / 64 * 64);
  const LevelDev D = dev_level(C);
  FX_DISPATCH_T(c->storage, {
    FX_DISPATCH_OP(C.op, {
      hipLaunchKernelGGL((k_coarse_cg<TV, OPV>), dim3(1), block, 0, s, D, (const TV*)C.rhs, (TV*)C.p, c->coarse_scale,
                         m, info);
    });
  });
  FX_LAUNCHED("k_coarse_cg");
  return FX_OK;
}

static int launch_direct(const fx_mg_config* c, hipStream_t s) {
  if (c->coarse_sweeps < 0) return launch_coarse_cg(c, nullptr, s);
  const fx_mg_level& C = c->level[c->levels - 1];
  int m = C.nx * C.ny * C.nz;
  FX_REQUIRE(m <= 4096, FX_EINVAL, "direct coarse solve limited to 4096 cells, got %d", m);
  dim3 grid((m + 3) / 4);
  FX_DISPATCH_T(c->storage, {
    hipLaunchKernelGGL((k_direct<TV>), grid, dim3(256), 0, s, m, c->coarse_matrix, (const TV*)C.rhs, (TV*)C.p,
                       c->coarse_scale);
  });
  FX_LAUNCHED("k_direct");
  return FX_OK;
}

template <typename T>
__global__ void k_zero(int64_t n, T* out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = from_f<T>(0.0f);
}

static int launch_fill_zero(fx_mg_config* c, void* ptr, int64_t n, hipStream_t s) {
  FX_DISPATCH_T(c->storage, {
    hipLaunchKernelGGL((k_zero<TV>), dim3((n + 255) / 256), dim3(256), 0, s, n, (TV*)ptr);
  });
  FX_LAUNCHED("k_zero");
  return FX_OK;
}

// Levels above this many cells keep separate launches: there the vector
// kernels run at full occupancy, and the halo recomputation of a fused leg
// (about 6 stencil evaluations per cell on 8^3 blocks) costs more than the
// launches it saves (measured at 128^3: 42 + 48 us fused against 22 + 25 us).
constexpr int64_t kLegMaxCells = int64_t(1) << 18;

// Can level l (not the coarsest) run its V-cycle legs fused?
static bool legs_fused(const fx_mg_config* c, int l, bool zero_guess) {
  const fx_mg_level& F = c->level[l];
  const fx_mg_level& C = c->level[l + 1];
  if (int64_t(F.nx) * F.ny * F.nz > kLegMaxCells) return false;
  // renormalised prolongation reads the coarse activity bytes (a box level is all active)
  const bool renorm_known = C.op == FX_OP_BOX || (op_has_edges(C.op) && C.active);
  return !c->separate_legs && l >= 1 && zero_guess && c->pre >= 1 && c->pre <= 3 && c->post >= 1 && c->post <= 3 &&
         c->restriction == FX_RESTRICT_MEAN &&
         (c->prolongation == FX_PROLONG_TRILINEAR || (c->prolongation == FX_PROLONG_RENORM && renorm_known)) &&
         (F.op == FX_OP_BOX || F.op == FX_OP_PACKED) && F.p != F.q;
}

static dim3 leg_grid(const fx_mg_level& L, int& tiles_x, int& tiles_y) {
  tiles_x = (L.nx + kFx - 1) / kFx;
  tiles_y = (L.ny + kFy - 1) / kFy;
  return dim3(tiles_x * tiles_y * ((L.nz + kFz - 1) / kFz));
}

// down leg: pre-sweeps from zero into the array the separate sweeps would
// leave them in (L.p after an even count, L.q after an odd one), restricted
// residual into the coarse rhs
static int launch_leg_down(fx_mg_config* c, int l, float inv_dt, hipStream_t s) {
  fx_mg_level& L = c->level[l];
  fx_mg_level& C = c->level[l + 1];
  int tx, ty;
  const dim3 grid = leg_grid(L, tx, ty);
  const LevelDev D = dev_level(L);
  void* dst = (c->pre % 2) ? L.q : L.p;
#define FX_LEG_DOWN(OPL, SW)                                                                                   \
  hipLaunchKernelGGL((k_leg_down<TV, OPL, SW>), grid, dim3(kLegThreads), 0, s, D, tx, ty, (const TV*)L.rhs,     \
                     (TV*)dst, (TV*)C.rhs, inv_dt, c->omega)
#define FX_LEG_DOWN_S(OPL)                                                                                     \
  do {                                                                                                         \
    if (c->pre == 1) FX_LEG_DOWN(OPL, 1);                                                                      \
    else if (c->pre == 2) FX_LEG_DOWN(OPL, 2);                                                                 \
    else FX_LEG_DOWN(OPL, 3);                                                                                  \
  } while (0)
  FX_DISPATCH_T(c->storage, {
    if (L.op == FX_OP_PACKED) FX_LEG_DOWN_S(FX_OP_PACKED);
    else FX_LEG_DOWN_S(FX_OP_BOX);
  });
#undef FX_LEG_DOWN_S
#undef FX_LEG_DOWN
  FX_LAUNCHED("k_leg_down");
  if (c->pre % 2) {
    L.q = L.p;
    L.p = dst;
  }
  return FX_OK;
}

// up leg: prolonged correction and post-sweeps from L.p into L.q, then p/q swap
static int launch_leg_up(fx_mg_config* c, int l, float inv_dt, hipStream_t s) {
  fx_mg_level& L = c->level[l];
  const fx_mg_level& C = c->level[l + 1];
  int tx, ty;
  const dim3 grid = leg_grid(L, tx, ty);
  const LevelDev DF = dev_level(L), DC = dev_level(C);
  const uint8_t* cact = op_has_edges(C.op) ? C.active : nullptr;
#define FX_LEG_UP(OPF, MODE, SW)                                                                               \
  hipLaunchKernelGGL((k_leg_up<TV, OPF, MODE, SW>), grid, dim3(kLegThreads), 0, s, DF, DC, cact, tx, ty,       \
                     (const TV*)L.p, (const TV*)C.p, (const TV*)L.rhs, (TV*)L.q, inv_dt, c->omega)
#define FX_LEG_UP_S(OPF, MODE)                                                                                 \
  do {                                                                                                         \
    if (c->post == 1) FX_LEG_UP(OPF, MODE, 1);                                                                 \
    else if (c->post == 2) FX_LEG_UP(OPF, MODE, 2);                                                            \
    else FX_LEG_UP(OPF, MODE, 3);                                                                              \
  } while (0)
  FX_DISPATCH_T(c->storage, {
    if (c->prolongation == FX_PROLONG_RENORM) {
      if (L.op == FX_OP_PACKED) FX_LEG_UP_S(FX_OP_PACKED, FX_PROLONG_RENORM);
      else FX_LEG_UP_S(FX_OP_BOX, FX_PROLONG_RENORM);
    } else {
      if (L.op == FX_OP_PACKED) FX_LEG_UP_S(FX_OP_PACKED, FX_PROLONG_TRILINEAR);
      else FX_LEG_UP_S(FX_OP_BOX, FX_PROLONG_TRILINEAR);
    }
  });
#undef FX_LEG_UP_S
#undef FX_LEG_UP
  FX_LAUNCHED("k_leg_up");
  void* t = L.p;
  L.p = L.q;
  L.q = t;
  return FX_OK;
}

// Recursive V-cycle.  zero_guess: the level's current iterate is 0.
int vcycle(fx_mg_config* c, int l, float inv_dt, bool zero_guess, hipStream_t s) {
  fx_mg_level& L = c->level[l];
  if (l == c->levels - 1) {
    if (c->coarse_sweeps > 0) return smooth_n(c, l, c->coarse_sweeps, inv_dt, zero_guess, s);
    return launch_direct(c, s);
  }
  if (legs_fused(c, l, zero_guess)) {
    if (int e = launch_leg_down(c, l, inv_dt, s)) return e;
    if (int e = vcycle(c, l + 1, 1.0f, true, s)) return e;
    return launch_leg_up(c, l, inv_dt, s);
  }
  if (c->pre > 0) {
    if (int e = smooth_n(c, l, c->pre, inv_dt, zero_guess, s)) return e;
  } else if (zero_guess) {
    if (int e = launch_fill_zero(c, L.p, int64_t(L.nx) * L.ny * L.nz, s)) return e;
  }
  if (int e = launch_residual_restrict(c, l, inv_dt, s)) return e;
  if (int e = vcycle(c, l + 1, 1.0f, true, s)) return e;
  if (int e = launch_prolong(c, l, L.p, c->level[l + 1].p, L.p, s)) return e;
  return smooth_n(c, l, c->post, inv_dt, false, s);
}

// `count` sweeps on level l from its current p (p/q swap per sweep)
int smooth_level(fx_mg_config* c, int l, int count, hipStream_t s) {
  return smooth_n(c, l, count, 1.0f, false, s);
}

fx_mg_level level0_of(fx_grid g, const uint8_t* flags) {
  fx_mg_level L{};
  L.nx = g.nx; L.ny = g.ny; L.nz = g.nz;
  L.op = flags ? FX_OP_FLAGS : FX_OP_BOX;
  float ix = g.nx / g.ex, iy = g.ny / g.ey, iz = g.nz / g.ez;
  L.w[0] = ix * ix; L.w[1] = iy * iy; L.w[2] = iz * iz;
  L.op_data = flags;
  return L;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_mg_vcycle(fx_mg_config* cfg, float inv_dt, void* stream) {
  if (int e = check_config(cfg)) return e;
  return vcycle(cfg, 0, inv_dt, false, S(stream));
}

int fx_mg_smooth(fx_mg_config* cfg, int32_t index, int32_t count, float inv_dt, void* stream) {
  if (int e = check_config(cfg)) return e;
  FX_REQUIRE(index >= 0 && index < cfg->levels && count >= 0, FX_EINVAL, "bad level or count");
  return smooth_n(cfg, index, count, inv_dt, false, S(stream));
}

int fx_mg_residual(const fx_mg_config* cfg, int32_t index, float inv_dt, void* out, void* stream) {
  if (int e = check_config(cfg)) return e;
  FX_REQUIRE(index >= 0 && index < cfg->levels && out, FX_EINVAL, "bad level or output");
  const fx_mg_level& L = cfg->level[index];
  LevelDev D = dev_level(L);
  FX_DISPATCH_T(cfg->storage, FX_DISPATCH_OP(L.op, {
    hipLaunchKernelGGL((k_residual<TV, OPV>), level_grid(L), kBlock, 0, S(stream), D, (const TV*)L.p,
                       (const TV*)L.rhs, (TV*)out, inv_dt);
  }));
  FX_LAUNCHED("fx_mg_residual");
  return FX_OK;
}

int fx_mg_residual_restrict(fx_mg_config* cfg, int32_t index, float inv_dt, void* stream) {
  if (int e = check_config(cfg)) return e;
  FX_REQUIRE(index >= 0 && index + 1 < cfg->levels, FX_EINVAL, "bad level");
  return launch_residual_restrict(cfg, index, inv_dt, S(stream));
}

int fx_mg_restrict(const fx_mg_config* cfg, int32_t index, const void* fine, void* coarse, void* stream) {
  if (int e = check_config(cfg)) return e;
  FX_REQUIRE(index >= 0 && index + 1 < cfg->levels && fine && coarse, FX_EINVAL, "bad level or arrays");
  const fx_mg_level& F = cfg->level[index];
  const fx_mg_level& C = cfg->level[index + 1];
  if (cfg->restriction == FX_RESTRICT_MEAN) {
    FX_DISPATCH_T(cfg->storage, {
      hipLaunchKernelGGL((k_restrict_mean<TV>), level_grid(C), kBlock, 0, S(stream), C.nx, C.ny, F.nx, F.ny,
                         (const TV*)fine, (TV*)coarse);
    });
  } else {
    // a plain restriction of any array: no coarse masking
    FX_DISPATCH_T(cfg->storage, {
      hipLaunchKernelGGL((k_restrict_full<TV, FX_OP_BOX>), level_grid(C), kBlock, 0, S(stream), dev_level(C), F.nx,
                         F.ny, F.nz, (const TV*)fine, (TV*)coarse);
    });
  }
  FX_LAUNCHED("fx_mg_restrict");
  return FX_OK;
}

int fx_mg_prolong(const fx_mg_config* cfg, int32_t index, const void* fine_in, const void* coarse, void* fine_out,
                  void* stream) {
  if (int e = check_config(cfg)) return e;
  FX_REQUIRE(index >= 0 && index + 1 < cfg->levels && coarse && fine_out, FX_EINVAL, "bad level or arrays");
  return launch_prolong(cfg, index, fine_in, coarse, fine_out, S(stream));
}

int fx_mg_direct(const fx_mg_config* cfg, void* stream) {
  if (int e = check_config(cfg)) return e;
  FX_REQUIRE(cfg->coarse_matrix || cfg->coarse_sweeps < 0, FX_EINVAL, "no coarse matrix");
  return launch_direct(cfg, S(stream));
}

int fx_mg_coarse_cg(const fx_mg_config* cfg, double* info, void* stream) {
  if (int e = check_config(cfg)) return e;
  return launch_coarse_cg(cfg, info, S(stream));
}

}  // extern "C"
