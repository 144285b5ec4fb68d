// Projection kernels: masks, walls, divergence, Jacobi, gradient.
// One thread per cell or face, lanes of a wavefront walk the contiguous x axis.
#include <cstdarg>
#include <cstdio>

#include "common.hpp"

namespace fx {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

int fail_hip(hipError_t e, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return FX_EDEVICE;
}

int check_grid(const fx_grid& g) {
  FX_REQUIRE(g.nx >= 1 && g.ny >= 1 && g.nz >= 1 && g.nx <= 4096 && g.ny <= 4096 && g.nz <= 4096, FX_EINVAL,
             "grid dimensions must be in [1, 4096], got (%d, %d, %d)", g.nx, g.ny, g.nz);
  FX_REQUIRE(g.ex > 0 && g.ey > 0 && g.ez > 0, FX_EINVAL, "grid extent must be positive");
  // kernels index cells and faces with 32-bit offsets
  FX_REQUIRE(int64_t(g.nx + 1) * (g.ny + 1) * (g.nz + 1) < (int64_t(1) << 31), FX_EINVAL,
             "grid too large: (nx+1)(ny+1)(nz+1) must stay below 2^31");
  return FX_OK;
}

static const dim3 kBlock(64, 4, 1);

// ---------------------------------------------------------------------------
__global__ void k_solid_raster(int nx, int ny, int nz, float hx, float hy, float hz, int kind, float r00,
                               float r01, float r02, float r03, float r10, float r11, float r12, float r13,
                               float r20, float r21, float r22, float r23, uint8_t* solid) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  float px = (x + 0.5f) * hx, py = (y + 0.5f) * hy, pz = (z + 0.5f) * hz;
  float qx = r00 * px + r01 * py + r02 * pz + r03;
  float qy = r10 * px + r11 * py + r12 * pz + r13;
  float qz = r20 * px + r21 * py + r22 * pz + r23;
  bool inside = kind == 0 ? (qx * qx + qy * qy + qz * qz <= 1.0f)
                          : (fabsf(qx) <= 1.0f && fabsf(qy) <= 1.0f && fabsf(qz) <= 1.0f);
  if (inside) solid[idx3(x, y, z, nx, ny)] = 1;
}

// open-face flags of the cells [x0, x1) x [y0, y1) x [z0 + blockIdx.z]
__global__ void k_build_flags(int nx, int ny, int nz, const uint8_t* __restrict__ solid, uint8_t* flags, int x0, int y0,
                              int z0, int x1, int y1) {
  int x = x0 + blockIdx.x * blockDim.x + threadIdx.x, y = y0 + blockIdx.y * blockDim.y + threadIdx.y,
      z = z0 + blockIdx.z;
  if (x >= x1 || y >= y1) return;
  int64_t i = idx3(x, y, z, nx, ny);
  uint8_t f = box_flags(x, y, z, nx, ny, nz);
  if (solid) {
    if (solid[i]) {
      f = 0;
    } else {
      int64_t sx = 1, sy = nx, sz = int64_t(nx) * ny;
      if ((f & FLAG_XLO) && solid[i - sx]) f &= ~FLAG_XLO;
      if ((f & FLAG_XHI) && solid[i + sx]) f &= ~FLAG_XHI;
      if ((f & FLAG_YLO) && solid[i - sy]) f &= ~FLAG_YLO;
      if ((f & FLAG_YHI) && solid[i + sy]) f &= ~FLAG_YHI;
      if ((f & FLAG_ZLO) && solid[i - sz]) f &= ~FLAG_ZLO;
      if ((f & FLAG_ZHI) && solid[i + sz]) f &= ~FLAG_ZHI;
    }
  }
  flags[i] = f;
}

__global__ void k_solid_clear(int64_t n, const uint8_t* __restrict__ solid, const float* __restrict__ in,
                              float* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = solid[i] ? 0.0f : in[i];
}

__global__ void k_fill(int64_t n, float v, float* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

__global__ void k_scale(int64_t n, float f, const float* __restrict__ in, float* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * f;
}

// Flags of the cell (x,y,z); box flags when no mask is given.
__device__ __forceinline__ uint8_t cell_flags(const uint8_t* flags, int x, int y, int z, int nx, int ny,
                                              int nz) {
  return flags ? flags[idx3(x, y, z, nx, ny)] : box_flags(x, y, z, nx, ny, nz);
}

// A face along axis a at face index (x,y,z) is open iff it is interior and both
// neighbouring cells are fluid: equivalently the high-side cell's low-face bit.
__device__ __forceinline__ bool face_open(const uint8_t* flags, int a, int x, int y, int z, int nx, int ny,
                                          int nz) {
  int c = a == 0 ? x : (a == 1 ? y : z);
  int n = a == 0 ? nx : (a == 1 ? ny : nz);
  if (c <= 0 || c >= n) return false;
  uint8_t f = cell_flags(flags, x, y, z, nx, ny, nz);
  return (f >> (2 * a)) & 1;
}

// Threads cover (nx+1, ny+1, nz+1); each writes the U, V, W face it owns.
// Closed faces read 0, or the wall velocity of moving colliders (bu/bv/bw,
// shaders/moving_boundary.glsl) when given.
__device__ __forceinline__ float closed_value(const float* __restrict__ b, int64_t i) { return b ? b[i] : 0.0f; }

__global__ void k_enforce_walls(int nx, int ny, int nz, const uint8_t* __restrict__ flags,
                                const float* __restrict__ u, const float* __restrict__ v,
                                const float* __restrict__ w, const float* __restrict__ bu,
                                const float* __restrict__ bv, const float* __restrict__ bw, float* __restrict__ uo,
                                float* __restrict__ vo, float* __restrict__ wo) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x > nx || y > ny) return;
  if (y < ny && z < nz) {
    int64_t i = idx3(x, y, z, nx + 1, ny);
    uo[i] = face_open(flags, 0, x, y, z, nx, ny, nz) ? u[i] : closed_value(bu, i);
  }
  if (x < nx && z < nz) {
    int64_t i = idx3(x, y, z, nx, ny + 1);
    vo[i] = face_open(flags, 1, x, y, z, nx, ny, nz) ? v[i] : closed_value(bv, i);
  }
  if (x < nx && y < ny) {
    int64_t i = idx3(x, y, z, nx, ny);
    wo[i] = face_open(flags, 2, x, y, z, nx, ny, nz) ? w[i] : closed_value(bw, i);
  }
}

__global__ void k_divergence(int nx, int ny, int nz, float ix, float iy, float iz,
                             const uint8_t* __restrict__ flags, const float* __restrict__ u,
                             const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ bu,
                             const float* __restrict__ bv, const float* __restrict__ bw, float* __restrict__ div) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  uint8_t f = cell_flags(flags, x, y, z, nx, ny, nz);
  float d = 0.0f;
  if (f & FLAG_FLUID) {
    const int64_t iu = idx3(x, y, z, nx + 1, ny), iv = idx3(x, y, z, nx, ny + 1), iw = idx3(x, y, z, nx, ny);
    const int64_t iv1 = iv + nx, iw1 = iw + int64_t(nx) * ny;
    float u0 = (f & FLAG_XLO) ? u[iu] : closed_value(bu, iu), u1 = (f & FLAG_XHI) ? u[iu + 1] : closed_value(bu, iu + 1);
    float v0 = (f & FLAG_YLO) ? v[iv] : closed_value(bv, iv), v1 = (f & FLAG_YHI) ? v[iv1] : closed_value(bv, iv1);
    float w0 = (f & FLAG_ZLO) ? w[iw] : closed_value(bw, iw), w1 = (f & FLAG_ZHI) ? w[iw1] : closed_value(bw, iw1);
    d = (u1 - u0) * ix + (v1 - v0) * iy + (w1 - w0) * iz;
  }
  div[idx3(x, y, z, nx, ny)] = d;
}

__global__ void k_jacobi(int nx, int ny, int nz, float wx, float wy, float wz, float inv_dt, float omega,
                         const uint8_t* __restrict__ flags, const float* __restrict__ p,
                         const float* __restrict__ rhs, float* __restrict__ out) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  int64_t i = idx3(x, y, z, nx, ny);
  uint8_t f = cell_flags(flags, x, y, z, nx, ny, nz);
  int64_t sy = nx, sz = int64_t(nx) * ny;
  const float c = p[i];
  float s = 0.0f, d = 0.0f;  // s = sum w (p_n - p_c)
  if (f & FLAG_XLO) { s += wx * (p[i - 1] - c); d += wx; }
  if (f & FLAG_XHI) { s += wx * (p[i + 1] - c); d += wx; }
  if (f & FLAG_YLO) { s += wy * (p[i - sy] - c); d += wy; }
  if (f & FLAG_YHI) { s += wy * (p[i + sy] - c); d += wy; }
  if (f & FLAG_ZLO) { s += wz * (p[i - sz] - c); d += wz; }
  if (f & FLAG_ZHI) { s += wz * (p[i + sz] - c); d += wz; }
  float r = 0.0f;
  if ((f & FLAG_FLUID) && d != 0.0f) r = c + omega * (s - rhs[i] * inv_dt) / d;
  out[i] = r;
}

__global__ void k_gradient(int nx, int ny, int nz, float ix, float iy, float iz, float scale,
                           const uint8_t* __restrict__ flags, const float* __restrict__ p,
                           const float* __restrict__ dp, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                           const float* __restrict__ bu, const float* __restrict__ bv, const float* __restrict__ bw,
                           float* __restrict__ uo, float* __restrict__ vo, float* __restrict__ wo) {
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x > nx || y > ny) return;
  int64_t sy = nx, sz = int64_t(nx) * ny;
  int64_t c = idx3(x, y, z, nx, ny);
  if (y < ny && z < nz) {
    int64_t i = idx3(x, y, z, nx + 1, ny);
    float r = closed_value(bu, i);
    if (face_open(flags, 0, x, y, z, nx, ny, nz)) {
      float g = p[c] - p[c - 1];
      if (dp) g += dp[c] - dp[c - 1];
      r = u[i] - scale * g * ix;
    }
    uo[i] = r;
  }
  if (x < nx && z < nz) {
    int64_t i = idx3(x, y, z, nx, ny + 1);
    float r = closed_value(bv, i);
    if (face_open(flags, 1, x, y, z, nx, ny, nz)) {
      float g = p[c] - p[c - sy];
      if (dp) g += dp[c] - dp[c - sy];
      r = v[i] - scale * g * iy;
    }
    vo[i] = r;
  }
  if (x < nx && y < ny) {
    float r = closed_value(bw, c);
    if (face_open(flags, 2, x, y, z, nx, ny, nz)) {
      float g = p[c] - p[c - sz];
      if (dp) g += dp[c] - dp[c - sz];
      r = w[c] - scale * g * iz;
    }
    wo[c] = r;
  }
}

// ---------------------------------------------------------------------------
// Vector forms of divergence and gradient (nx % 4 == 0, 16-byte aligned cell,
// V, W and flag arrays): lane q owns the four cells x0..x0+3 of one row, reads
// cells, V and W rows and the flag word with 16-byte / 4-byte loads (U rows
// are nx+1 long, so U goes through scalar loads), and blocks are spread so
// each XCD sweeps a contiguous range of rows, keeping the y/z neighbour rows
// in its own L2.  Same operations per face/cell as the one-cell kernels.
__device__ __forceinline__ int xcd_row_tile(int b, int nb) {
  return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

__device__ __forceinline__ float4 ld4f(const float* __restrict__ p, int i) {
  return *reinterpret_cast<const float4*>(p + i);
}
__device__ __forceinline__ void st4f(float* __restrict__ p, int i, float4 v) {
  *reinterpret_cast<float4*>(p + i) = v;
}
__device__ __forceinline__ float comp(const float4& v, int k) {
  return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

__device__ __forceinline__ uint32_t row_flags(const uint8_t* __restrict__ flags, int x0, int y, int z, int nx, int ny,
                                              int nz, int c) {
  if (flags) return *reinterpret_cast<const uint32_t*>(flags + c);
  uint32_t w = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) w |= uint32_t(box_flags(x0 + k, y, z, nx, ny, nz)) << (8 * k);
  return w;
}

__device__ __forceinline__ float bound_abs(float v) { return isfinite(v) ? fabsf(v) : 3.4e38f; }

// bu/bv/bw (nullable): the wall velocity closed faces read (moving colliders).
// part (nullable): the block's max |U|, |V|, |W| over the face values its
// lanes load, which over all lanes are every face of the three arrays (the
// lanes' U rows end at the next lane's first face, the V and W loads take
// the faces above the lane's cells as well), at part[axis * blocks + block]
// (non-finite values count as 3.4e38, as in fx_max_abs).
__global__ __launch_bounds__(256) void k_divergence_q(int nx, int ny, int nz, float ix, float iy, float iz,
                                                      const uint8_t* __restrict__ flags, const float* __restrict__ u,
                                                      const float* __restrict__ v, const float* __restrict__ w,
                                                      const float* __restrict__ bu, const float* __restrict__ bv,
                                                      const float* __restrict__ bw, float* __restrict__ div, int nq,
                                                      float* __restrict__ part) {
  const int tile = xcd_row_tile(blockIdx.x, gridDim.x);
  const int q = tile * 256 + threadIdx.x;
  float mu = 0.0f, mv = 0.0f, mw = 0.0f;
  if (q < nq) {
    const int qx = nx >> 2, row = q / qx, x0 = (q - row * qx) << 2, z = row / ny, y = row - z * ny;
    const int c = row * nx + x0, iu = row * (nx + 1) + x0, iv = (y + (ny + 1) * z) * nx + x0, sz = nx * ny;
    const uint32_t fw = row_flags(flags, x0, y, z, nx, ny, nz, c);
    float uu[5];
  #pragma unroll
    for (int k = 0; k < 5; ++k) uu[k] = u[iu + k];
    const float4 v0 = ld4f(v, iv), v1 = ld4f(v, iv + nx), w0 = ld4f(w, c), w1 = ld4f(w, c + sz);
    float d[4];
  #pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint8_t f = uint8_t(fw >> (8 * k));
      d[k] = 0.0f;
      if (f & FLAG_FLUID) {
        const float a0 = (f & FLAG_XLO) ? uu[k] : closed_value(bu, iu + k),
                    a1 = (f & FLAG_XHI) ? uu[k + 1] : closed_value(bu, iu + k + 1);
        const float b0 = (f & FLAG_YLO) ? comp(v0, k) : closed_value(bv, iv + k),
                    b1 = (f & FLAG_YHI) ? comp(v1, k) : closed_value(bv, iv + nx + k);
        const float c0 = (f & FLAG_ZLO) ? comp(w0, k) : closed_value(bw, c + k),
                    c1 = (f & FLAG_ZHI) ? comp(w1, k) : closed_value(bw, c + sz + k);
        d[k] = (a1 - a0) * ix + (b1 - b0) * iy + (c1 - c0) * iz;
      }
    }
    st4f(div, c, make_float4(d[0], d[1], d[2], d[3]));
    if (part) {
  #pragma unroll
      for (int k = 0; k < 5; ++k) mu = fmaxf(mu, bound_abs(uu[k]));
  #pragma unroll
      for (int k = 0; k < 4; ++k) {
        mv = fmaxf(mv, fmaxf(bound_abs(comp(v0, k)), bound_abs(comp(v1, k))));
        mw = fmaxf(mw, fmaxf(bound_abs(comp(w0, k)), bound_abs(comp(w1, k))));
      }
    }
  }
  if (part) {  // uniform over the grid
    __shared__ float m[3][4];
    mu = wave_max(mu);
    mv = wave_max(mv);
    mw = wave_max(mw);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
      m[0][wv] = mu;
      m[1][wv] = mv;
      m[2][wv] = mw;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      const int a = threadIdx.x;
      part[a * gridDim.x + blockIdx.x] = fmaxf(fmaxf(m[a][0], m[a][1]), fmaxf(m[a][2], m[a][3]));
    }
  }
}

// vmax[axis] = max over the `blocks` block partials of that axis (one block of
// 1024 threads per axis, four loads in flight per thread)
__global__ __launch_bounds__(1024) void k_bound_finish(const float* __restrict__ part, int blocks,
                                                       float* __restrict__ vmax) {
  __shared__ float m[16];
  const float* p = part + int64_t(blockIdx.x) * blocks;
  float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
  int b = threadIdx.x;
  for (; b + 3 * 1024 < blocks; b += 4 * 1024) {
    r0 = fmaxf(r0, p[b]);
    r1 = fmaxf(r1, p[b + 1024]);
    r2 = fmaxf(r2, p[b + 2048]);
    r3 = fmaxf(r3, p[b + 3072]);
  }
  for (; b < blocks; b += 1024) r0 = fmaxf(r0, p[b]);
  float r = wave_max(fmaxf(fmaxf(r0, r1), fmaxf(r2, r3)));
  if ((threadIdx.x & 63) == 0) m[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 16; ++k) r = fmaxf(r, m[k]);
    vmax[blockIdx.x] = fmaxf(r, m[0]);
  }
}

// Gradient correction of the faces at the low sides of the lane's cells (and
// the closed faces at x = nx, y = ny, z = nz next to the last cells); with dp
// the neighbour differences of p and dp are formed separately, and psum (if
// given) receives p + dp.  Closed faces take 0, or with bu/bv/bw the wall
// velocity of moving colliders.
__global__ __launch_bounds__(256) void k_gradient_q(int nx, int ny, int nz, float ix, float iy, float iz, float scale,
                                                    const uint8_t* __restrict__ flags, const float* __restrict__ p,
                                                    const float* __restrict__ dp, const float* __restrict__ u,
                                                    const float* __restrict__ v, const float* __restrict__ w,
                                                    const float* __restrict__ bu, const float* __restrict__ bv,
                                                    const float* __restrict__ bw, float* __restrict__ uo,
                                                    float* __restrict__ vo, float* __restrict__ wo,
                                                    float* __restrict__ psum, int nq) {
  const int q = xcd_row_tile(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (q >= nq) return;
  const int qx = nx >> 2, row = q / qx, x0 = (q - row * qx) << 2, z = row / ny, y = row - z * ny;
  const int c = row * nx + x0, iu = row * (nx + 1) + x0, iv = (y + (ny + 1) * z) * nx + x0, sz = nx * ny;
  const uint32_t fw = row_flags(flags, x0, y, z, nx, ny, nz, c);
  // unconditional loads (the own cell stands in for a missing lower
  // neighbour: its face is closed and never reads it), so one wait covers all
  const int cxm = x0 > 0 ? c - 1 : c, cym = y > 0 ? c - nx : c, czm = z > 0 ? c - sz : c;
  const float4 pc = ld4f(p, c), pym = ld4f(p, cym), pzm = ld4f(p, czm);
  const float pxm = p[cxm];
  float4 dc = make_float4(0.f, 0.f, 0.f, 0.f), dym = dc, dzm = dc;
  float dxm = 0.0f;
  if (dp) {
    dc = ld4f(dp, c);
    dxm = dp[cxm];
    dym = ld4f(dp, cym);
    dzm = ld4f(dp, czm);
  }
  float uu[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) uu[k] = u[iu + k];
  const float4 vv = ld4f(v, iv), ww = ld4f(w, c);
  float ur[4], vr[4], wr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint8_t f = uint8_t(fw >> (8 * k));
    const float pk = comp(pc, k);
    // x face at x0 + k: open iff interior and the cell's low-x bit is set
    if (x0 + k > 0 && (f & FLAG_XLO)) {
      float g = pk - (k == 0 ? pxm : comp(pc, k - 1));
      if (dp) g += comp(dc, k) - (k == 0 ? dxm : comp(dc, k - 1));
      ur[k] = uu[k] - scale * g * ix;
    } else {
      ur[k] = closed_value(bu, iu + k);
    }
    if (y > 0 && (f & FLAG_YLO)) {
      float g = pk - comp(pym, k);
      if (dp) g += comp(dc, k) - comp(dym, k);
      vr[k] = comp(vv, k) - scale * g * iy;
    } else {
      vr[k] = closed_value(bv, iv + k);
    }
    if (z > 0 && (f & FLAG_ZLO)) {
      float g = pk - comp(pzm, k);
      if (dp) g += comp(dc, k) - comp(dzm, k);
      wr[k] = comp(ww, k) - scale * g * iz;
    } else {
      wr[k] = closed_value(bw, c + k);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) uo[iu + k] = ur[k];
  if (x0 + 4 == nx) uo[iu + 4] = closed_value(bu, iu + 4);
  st4f(vo, iv, make_float4(vr[0], vr[1], vr[2], vr[3]));
  if (y == ny - 1)
    st4f(vo, iv + nx, make_float4(closed_value(bv, iv + nx), closed_value(bv, iv + nx + 1),
                                  closed_value(bv, iv + nx + 2), closed_value(bv, iv + nx + 3)));
  st4f(wo, c, make_float4(wr[0], wr[1], wr[2], wr[3]));
  if (z == nz - 1)
    st4f(wo, c + sz, make_float4(closed_value(bw, c + sz), closed_value(bw, c + sz + 1), closed_value(bw, c + sz + 2),
                                 closed_value(bw, c + sz + 3)));
  if (psum) st4f(psum, c, make_float4(pc.x + dc.x, pc.y + dc.y, pc.z + dc.z, pc.w + dc.w));
}

__global__ void k_combine(int64_t n, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// Initial fields of reset(): kind 0 = scalar scale*w^2 of the soft sphere
// (shaders/seed.glsl, temperature_seed.glsl); kind 1..3 = face component a of
// base + rotation about the vertical axis (shaders/mac_seed.glsl).
__global__ void k_seed(int nx, int ny, int nz, float ex, float ey, float ez, int kind, float cx, float cy, float cz,
                       float radius, float scale, float bx, float by, float bz, float omega, float* out) {
  int a = kind - 1;
  int sx = nx + (a == 0), sy = ny + (a == 1), sz = nz + (a == 2);
  int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= sx || y >= sy || z >= sz) return;
  if (kind == 0) {
    float px = (x + 0.5f) * ex / nx, py = (y + 0.5f) * ey / ny, pz = (z + 0.5f) * ez / nz;
    float dx = px - cx, dy = py - cy, dz = pz - cz;
    float w = fmaxf(1.0f - (dx * dx + dy * dy + dz * dz) / (radius * radius), 0.0f);
    out[idx3(x, y, z, nx, ny)] = scale * w * w;
    return;
  }
  float px = (x + (a == 0 ? 0.0f : 0.5f)) * ex / nx, py = (y + (a == 1 ? 0.0f : 0.5f)) * ey / ny;
  float qx = px - 0.5f * ex, qy = py - 0.5f * ey;
  float v = a == 0 ? bx - omega * qy : (a == 1 ? by + omega * qx : bz);
  out[idx3(x, y, z, sx, sy)] = v;
}

// host: can the row-vector divergence/gradient kernels run?
static bool rows_vec_ok(const fx_grid& g, const void* a, const void* b, const void* c, const void* d,
                        const void* e = nullptr, const void* f = nullptr) {
  auto al = [](const void* ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15u) == 0; };
  return g.nx % 4 == 0 && al(a) && al(b) && al(c) && al(d) && al(e) && al(f);
}

// open-face flags of the cells in box b = [x0, x1) x [y0, y1) x [z0, z1)
int launch_flags_box(const fx_grid& g, const int* b, const uint8_t* solid, uint8_t* flags, hipStream_t s) {
  const int bx = b[3] - b[0], by = b[4] - b[1], bz = b[5] - b[2];
  if (bx <= 0 || by <= 0 || bz <= 0) return FX_OK;
  hipLaunchKernelGGL(k_build_flags, dim3((bx + 63) / 64, (by + 3) / 4, bz), kBlock, 0, s, g.nx, g.ny, g.nz, solid,
                     flags, b[0], b[1], b[2], b[3], b[4]);
  FX_LAUNCHED("flags box");
  return FX_OK;
}

static int launch_gradient(const fx_grid& g, const uint8_t* flags, const float* p, const float* dp, const float* u,
                           const float* v, const float* w, float* uo, float* vo, float* wo, float* psum,
                           float scale, hipStream_t s, const char* name, const float* const* walls = nullptr) {
  const float ix = g.nx / g.ex, iy = g.ny / g.ey, iz = g.nz / g.ez;
  const float* bu = walls ? walls[0] : nullptr;
  const float* bv = walls ? walls[1] : nullptr;
  const float* bw = walls ? walls[2] : nullptr;
  if (rows_vec_ok(g, flags, p, dp, v, w, vo) && rows_vec_ok(g, wo, psum, nullptr, nullptr)) {
    const int nq = g.nx / 4 * g.ny * g.nz;
    hipLaunchKernelGGL(k_gradient_q, dim3((nq + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz, ix, iy, iz, scale,
                       flags, p, dp, u, v, w, bu, bv, bw, uo, vo, wo, psum, nq);
  } else {
    dim3 gr = grid3(g.nx + 1, g.ny + 1, g.nz + 1, kBlock);
    hipLaunchKernelGGL(k_gradient, gr, kBlock, 0, s, g.nx, g.ny, g.nz, ix, iy, iz, scale, flags, p, dp, u, v, w, bu,
                       bv, bw, uo, vo, wo);
    if (psum) {
      const int64_t n = int64_t(g.nx) * g.ny * g.nz;
      hipLaunchKernelGGL(k_combine, dim3((n + 255) / 256), dim3(256), 0, s, n, p, dp, psum);
    }
  }
  FX_LAUNCHED(name);
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" {

const char* fx_last_error(void) { return g_error; }
int fx_version(void) { return 1; }

int fx_solid_raster(fx_grid g, int32_t kind, const float* r, uint8_t* solid, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(kind == 0 || kind == 1, FX_EINVAL, "collider kind must be 0 (sphere) or 1 (box)");
  FX_REQUIRE(r && solid, FX_EINVAL, "null pointer");
  dim3 gr = grid3(g.nx, g.ny, g.nz, kBlock);
  hipLaunchKernelGGL(k_solid_raster, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, g.ex / g.nx, g.ey / g.ny,
                     g.ez / g.nz, kind, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11],
                     solid);
  FX_LAUNCHED("fx_solid_raster");
  return FX_OK;
}

int fx_seed(fx_grid g, int32_t kind, const float* params9, float* out, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(kind >= 0 && kind <= 3 && params9 && out, FX_EINVAL, "seed: bad arguments");
  const float* q = params9;
  dim3 gr = grid3(g.nx + 1, g.ny + 1, g.nz + 1, kBlock);
  hipLaunchKernelGGL(k_seed, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, g.ex, g.ey, g.ez, kind, q[0], q[1], q[2],
                     q[3], q[4], q[5], q[6], q[7], q[8], out);
  FX_LAUNCHED("fx_seed");
  return FX_OK;
}

int fx_build_flags(fx_grid g, const uint8_t* solid, uint8_t* flags, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(flags, FX_EINVAL, "null flags");
  dim3 gr = grid3(g.nx, g.ny, g.nz, kBlock);
  hipLaunchKernelGGL(k_build_flags, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, solid, flags, 0, 0, 0, g.nx, g.ny);
  FX_LAUNCHED("fx_build_flags");
  return FX_OK;
}

int fx_solid_clear(fx_grid g, const uint8_t* solid, const float* in, float* out, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(solid && in && out && in != out, FX_EINVAL, "solid_clear needs distinct input and output");
  int64_t n = int64_t(g.nx) * g.ny * g.nz;
  hipLaunchKernelGGL(k_solid_clear, dim3((n + 255) / 256), dim3(256), 0, S(stream), n, solid, in, out);
  FX_LAUNCHED("fx_solid_clear");
  return FX_OK;
}

int fx_fill(float* out, int64_t n, float value, void* stream) {
  FX_REQUIRE(out && n >= 0, FX_EINVAL, "fill: bad arguments");
  if (n == 0) return FX_OK;
  hipLaunchKernelGGL(k_fill, dim3((n + 255) / 256), dim3(256), 0, S(stream), n, value, out);
  FX_LAUNCHED("fx_fill");
  return FX_OK;
}

int fx_scale(const float* in, float* out, int64_t n, float factor, void* stream) {
  FX_REQUIRE(in && out && n >= 0, FX_EINVAL, "scale: bad arguments");
  if (n == 0) return FX_OK;
  hipLaunchKernelGGL(k_scale, dim3((n + 255) / 256), dim3(256), 0, S(stream), n, factor, in, out);
  FX_LAUNCHED("fx_scale");
  return FX_OK;
}

int fx_enforce_walls(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w, float* uo,
                     float* vo, float* wo, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(u && v && w && uo && vo && wo, FX_EINVAL, "null face array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  dim3 gr = grid3(g.nx + 1, g.ny + 1, g.nz + 1, kBlock);
  hipLaunchKernelGGL(k_enforce_walls, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, flags, u, v, w,
                     (const float*)nullptr, (const float*)nullptr, (const float*)nullptr, uo, vo, wo);
  FX_LAUNCHED("fx_enforce_walls");
  return FX_OK;
}

int fx_enforce_walls_moving(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                            const float* bu, const float* bv, const float* bw, float* uo, float* vo, float* wo,
                            void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(flags && u && v && w && bu && bv && bw && uo && vo && wo, FX_EINVAL, "null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  dim3 gr = grid3(g.nx + 1, g.ny + 1, g.nz + 1, kBlock);
  hipLaunchKernelGGL(k_enforce_walls, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, flags, u, v, w, bu, bv, bw, uo, vo,
                     wo);
  FX_LAUNCHED("fx_enforce_walls_moving");
  return FX_OK;
}

int fx_divergence(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w, float* div,
                  void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(u && v && w && div, FX_EINVAL, "null array");
  if (rows_vec_ok(g, flags, v, w, div)) {
    const int nq = g.nx / 4 * g.ny * g.nz;
    hipLaunchKernelGGL(k_divergence_q, dim3((nq + 255) / 256), dim3(256), 0, S(stream), g.nx, g.ny, g.nz,
                       g.nx / g.ex, g.ny / g.ey, g.nz / g.ez, flags, u, v, w, (const float*)nullptr,
                       (const float*)nullptr, (const float*)nullptr, div, nq, (float*)nullptr);
  } else {
    dim3 gr = grid3(g.nx, g.ny, g.nz, kBlock);
    hipLaunchKernelGGL(k_divergence, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, g.nx / g.ex, g.ny / g.ey,
                       g.nz / g.ez, flags, u, v, w, (const float*)nullptr, (const float*)nullptr,
                       (const float*)nullptr, div);
  }
  FX_LAUNCHED("fx_divergence");
  return FX_OK;
}

int fx_divergence_bound(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                        const float* bu, const float* bv, const float* bw, float* div, float* partials, float* vmax,
                        void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(u && v && w && div && partials && vmax, FX_EINVAL, "null array");
  FX_REQUIRE((bu && bv && bw) || (!bu && !bv && !bw), FX_EINVAL, "divergence bound: all or none of the walls");
  FX_REQUIRE(!bu || flags, FX_EINVAL, "divergence bound: walls need the flags");
  hipStream_t s = S(stream);
  if (rows_vec_ok(g, flags, v, w, div)) {
    const int nq = g.nx / 4 * g.ny * g.nz, blocks = (nq + 255) / 256;
    hipLaunchKernelGGL(k_divergence_q, dim3(blocks), dim3(256), 0, s, g.nx, g.ny, g.nz, g.nx / g.ex, g.ny / g.ey,
                       g.nz / g.ez, flags, u, v, w, bu, bv, bw, div, nq, partials);
    hipLaunchKernelGGL(k_bound_finish, dim3(3), dim3(1024), 0, s, (const float*)partials, blocks, vmax);
    FX_LAUNCHED("fx_divergence_bound");
    return FX_OK;
  }
  // one-cell form: the divergence, then the bound from the arrays
  dim3 gr = grid3(g.nx, g.ny, g.nz, kBlock);
  hipLaunchKernelGGL(k_divergence, gr, kBlock, 0, s, g.nx, g.ny, g.nz, g.nx / g.ex, g.ny / g.ey, g.nz / g.ez, flags,
                     u, v, w, bu, bv, bw, div);
  FX_LAUNCHED("fx_divergence_bound (one-cell)");
  const float* arrays[3] = {u, v, w};
  const int64_t sizes[3] = {int64_t(g.nx + 1) * g.ny * g.nz, int64_t(g.nx) * (g.ny + 1) * g.nz,
                            int64_t(g.nx) * g.ny * (g.nz + 1)};
  return fx_max_abs(3, arrays, sizes, vmax, stream);
}

int fx_divergence_moving(fx_grid g, const uint8_t* flags, const float* u, const float* v, const float* w,
                         const float* bu, const float* bv, const float* bw, float* div, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(flags && u && v && w && bu && bv && bw && div, FX_EINVAL, "null array");
  if (rows_vec_ok(g, flags, v, w, div)) {
    const int nq = g.nx / 4 * g.ny * g.nz;
    hipLaunchKernelGGL(k_divergence_q, dim3((nq + 255) / 256), dim3(256), 0, S(stream), g.nx, g.ny, g.nz,
                       g.nx / g.ex, g.ny / g.ey, g.nz / g.ez, flags, u, v, w, bu, bv, bw, div, nq, (float*)nullptr);
  } else {
    dim3 gr = grid3(g.nx, g.ny, g.nz, kBlock);
    hipLaunchKernelGGL(k_divergence, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, g.nx / g.ex, g.ny / g.ey,
                       g.nz / g.ez, flags, u, v, w, bu, bv, bw, div);
  }
  FX_LAUNCHED("fx_divergence_moving");
  return FX_OK;
}

int fx_jacobi_sweep(fx_grid g, const uint8_t* flags, const float* p, const float* rhs, float* out, float inv_dt,
                    float omega, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(p && rhs && out && out != p && out != rhs, FX_EINVAL,
             "Advection input/output textures must be distinct");
  float ix = g.nx / g.ex, iy = g.ny / g.ey, iz = g.nz / g.ez;
  dim3 gr = grid3(g.nx, g.ny, g.nz, kBlock);
  hipLaunchKernelGGL(k_jacobi, gr, kBlock, 0, S(stream), g.nx, g.ny, g.nz, ix * ix, iy * iy, iz * iz, inv_dt,
                     omega, flags, p, rhs, out);
  FX_LAUNCHED("fx_jacobi_sweep");
  return FX_OK;
}

int fx_gradient(fx_grid g, const uint8_t* flags, const float* p, const float* u, const float* v, const float* w,
                float* uo, float* vo, float* wo, float scale, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(p && u && v && w && uo && vo && wo, FX_EINVAL, "null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  return launch_gradient(g, flags, p, nullptr, u, v, w, uo, vo, wo, nullptr, scale, S(stream), "fx_gradient");
}

int fx_gradient_split(fx_grid g, const uint8_t* flags, const float* p, const float* dp, const float* u,
                      const float* v, const float* w, float* uo, float* vo, float* wo, float* psum, float scale,
                      void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(p && dp && u && v && w && uo && vo && wo, FX_EINVAL, "null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  FX_REQUIRE(psum != p && psum != dp, FX_EINVAL, "gradient: the pressure sum must not alias p or dp");
  return launch_gradient(g, flags, p, dp, u, v, w, uo, vo, wo, psum, scale, S(stream), "fx_gradient_split");
}

int fx_gradient_moving(fx_grid g, const uint8_t* flags, const float* p, const float* u, const float* v,
                       const float* w, const float* bu, const float* bv, const float* bw, float* uo, float* vo,
                       float* wo, float scale, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(flags && p && u && v && w && bu && bv && bw && uo && vo && wo, FX_EINVAL, "null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  const float* walls[3] = {bu, bv, bw};
  return launch_gradient(g, flags, p, nullptr, u, v, w, uo, vo, wo, nullptr, scale, S(stream), "fx_gradient_moving",
                         walls);
}

int fx_gradient_split_moving(fx_grid g, const uint8_t* flags, const float* p, const float* dp, const float* u,
                             const float* v, const float* w, const float* bu, const float* bv, const float* bw,
                             float* uo, float* vo, float* wo, float* psum, float scale, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(flags && p && dp && u && v && w && bu && bv && bw && uo && vo && wo, FX_EINVAL, "null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  FX_REQUIRE(psum != p && psum != dp, FX_EINVAL, "gradient: the pressure sum must not alias p or dp");
  const float* walls[3] = {bu, bv, bw};
  return launch_gradient(g, flags, p, dp, u, v, w, uo, vo, wo, psum, scale, S(stream), "fx_gradient_split_moving",
                         walls);
}

int fx_combine(const float* a, const float* b, float* out, int64_t n, void* stream) {
  FX_REQUIRE(a && b && out && n >= 0, FX_EINVAL, "combine: bad arrays");
  if (n == 0) return FX_OK;
  hipLaunchKernelGGL(k_combine, dim3((n + 255) / 256), dim3(256), 0, S(stream), n, a, b, out);
  FX_LAUNCHED("fx_combine");
  return FX_OK;
}

}  // extern "C"
