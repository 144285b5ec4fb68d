// Tolerance-driven solvers on top of the multigrid engine (no reference
// counterpart; the reference runs fixed cycle budgets):
//   * mixed-precision iterative refinement: FP32 iterate and residual, V-cycle
//     on the error equation in 16-bit storage with FP32 arithmetic;
//   * flexible preconditioned CG (Polak-Ribiere beta) with a V-cycle
//     preconditioner in any storage type.
// Norms and dot products accumulate in FP64 per block and are folded by one
// finalize block; alpha/beta are formed on the device.  The host reads one
// residual norm per iteration from `check_from` on.
#include <algorithm>

#include "quad.hpp"

namespace fx {

int resnorm_q(const fx_mg_level& L0, const float* x, const float* b, double* ws, int qr, int qb, hipStream_t s);
int norm2_q(int64_t n, const float* a, double* ws, int q, hipStream_t s);
int check_params(const fx_solve_params* p);
int unet_forward(const fx_unet* net, const float* r, const uint8_t* solid, double* ws, float post_scale,
                 int output_rms, float* z, hipStream_t s);
int smooth_level(fx_mg_config* c, int l, int count, hipStream_t s);

// 7-point operator of (x + s*e) at cell i, x fp32, e in storage T (or null).
template <int OP, typename T>
__device__ __forceinline__ float apply_shifted(const LevelDev& L, int X, int Y, int Z, int64_t i,
                                               const float* __restrict__ x, const T* __restrict__ e, float s,
                                               bool& active, float& center) {
  const int64_t sy = L.nx, sz = int64_t(L.nx) * L.ny;
  auto val = [&](int64_t j) { return x[j] + (e ? s * to_f(e[j]) : 0.0f); };
  center = val(i);
  float acc = 0.0f;
  if (op_has_edges(OP)) {
    float4 c = edge_at<OP>(L, i);
    active = c.w > 0.0f;
    float wxl = X > 0 ? edge_x<OP>(L, i - 1) : 0.0f, wyl = Y > 0 ? edge_y<OP>(L, i - sy) : 0.0f,
          wzl = Z > 0 ? edge_z<OP>(L, i - sz) : 0.0f;
    if (c.x != 0.0f) acc += c.x * (val(i + 1) - center);
    if (wxl != 0.0f) acc += wxl * (val(i - 1) - center);
    if (c.y != 0.0f) acc += c.y * (val(i + sy) - center);
    if (wyl != 0.0f) acc += wyl * (val(i - sy) - center);
    if (c.z != 0.0f) acc += c.z * (val(i + sz) - center);
    if (wzl != 0.0f) acc += wzl * (val(i - sz) - center);
  } else {
    uint8_t f = OP == FX_OP_FLAGS ? reinterpret_cast<const uint8_t*>(L.op)[i] : box_flags(X, Y, Z, L.nx, L.ny, L.nz);
    active = f & FLAG_FLUID;
    if (f & FLAG_XLO) acc += L.wx * (val(i - 1) - center);
    if (f & FLAG_XHI) acc += L.wx * (val(i + 1) - center);
    if (f & FLAG_YLO) acc += L.wy * (val(i - sy) - center);
    if (f & FLAG_YHI) acc += L.wy * (val(i + sy) - center);
    if (f & FLAG_ZLO) acc += L.wz * (val(i - sz) - center);
    if (f & FLAG_ZHI) acc += L.wz * (val(i + sz) - center);
  }
  return acc;
}

#define FX_FOR_CELLS_N(n)                                                                     \
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < (n);                  \
       i += int64_t(gridDim.x) * blockDim.x)

// x' = x + s_e*e (e may be null), r = b - L x', rT = r*inv_s, ws[q] partials of ||r||^2
template <int OP, typename T>
__global__ void k_ir_update(LevelDev L, const float* __restrict__ x, const T* __restrict__ e, float s_e,
                            const float* __restrict__ b, float* __restrict__ xo, T* __restrict__ rT, float inv_s,
                            double* ws, int q) {
  double rr = 0.0;
  FX_FOR_ROW_CELLS(L, X, Y, Z, i) {
    bool active;
    float center;
    float lap = apply_shifted<OP, T>(L, X, Y, Z, i, x, e, s_e, active, center);
    float r = active ? b[i] - lap : 0.0f;
    if (xo) xo[i] = active ? center : 0.0f;
    rT[i] = from_f<T>(r * inv_s);
    rr += double(r) * double(r);
  }
  block_partial(rr, ws, q);
}

template <typename T>
__global__ void k_rescale(int64_t n, T* __restrict__ a, float f) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = from_f<T>(to_f(a[i]) * f);
}

// ---------------------------------------------------------------------------
// Krylov kernels on A = -L (positive semi-definite on the fluid), c = -b
// ---------------------------------------------------------------------------
// r = L x - b, rT = -r*inv_s, partial ||r||^2 -> q
template <int OP, typename T>
__global__ void k_pcg_init(LevelDev L, const float* __restrict__ x, const float* __restrict__ b, float* __restrict__ r,
                           T* __restrict__ rT, float inv_s, double* ws, int q) {
  double rr = 0.0;
  FX_FOR_ROW_CELLS(L, X, Y, Z, i) {
    bool active;
    float center;
    float lap = apply_shifted<OP, float>(L, X, Y, Z, i, x, (const float*)nullptr, 0.0f, active, center);
    float rv = active ? lap - b[i] : 0.0f;
    r[i] = rv;
    rT[i] = from_f<T>(-rv * inv_s);
    rr += double(rv) * double(rv);
  }
  block_partial(rr, ws, q);
}

// Ap = -L p ; partial p.Ap -> q
template <int OP>
__global__ void k_pcg_apply(LevelDev L, const float* __restrict__ p, float* __restrict__ Ap, double* ws, int q) {
  double acc = 0.0;
  FX_FOR_ROW_CELLS(L, X, Y, Z, i) {
    bool active;
    float center;
    float lap = apply_shifted<OP, float>(L, X, Y, Z, i, p, (const float*)nullptr, 0.0f, active, center);
    float a = active ? -lap : 0.0f;
    Ap[i] = a;
    acc += double(center) * double(a);
  }
  block_partial(acc, ws, q);
}

// alpha = ws[q_rz]/ws[q_pap]; x += alpha p; r -= alpha Ap; rT = -r*inv_s;
// Scale s of the 16-bit preconditioner residual (rT = -r/s, z_A = -s e).  With
// slot < 0 the host's values are used as they are (scale 1 for fp32/bf16);
// otherwise (fp16) s lives in ws[slot] and the update kernel derives the next
// iteration's scale, max(sqrt(ws[0]) * rms_div * 0.3, 1e-30), from the
// ||r||^2 finalised after the previous update (the 0.3 anticipates this
// update's reduction) and leaves it in ws[next]: FP16 then follows the same
// host residual-check schedule as the other storage types.
struct ScaleRef {
  float zscale;  // -s (host value)
  float inv_s;   // 1/s of the rT this kernel writes (host value)
  int slot;      // ws slot of the current s, or -1
  int next;      // ws slot for the next s (update kernel)
  double rms_div;
  __device__ __forceinline__ float z(const double* ws) const { return slot < 0 ? zscale : float(-ws[slot]); }
  // 1/s_next for rT; block 0 stores s_next
  __device__ __forceinline__ float inv_next(double* ws) const {
    if (slot < 0) return inv_s;
    const double s_next = fmax(sqrt(ws[0]) * rms_div * 0.3, 1e-30);
    if (blockIdx.x == 0 && threadIdx.x == 0) ws[next] = s_next;
    return float(1.0 / s_next);
  }
};

__global__ void k_set_ws(double* ws, int q, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ws[q] = v;
}

// partials ||r||^2 -> q_rr, r.(zold*zscale) -> q_cross
template <typename T>
__global__ void k_pcg_update(int64_t n, double* ws, int q_rz, int q_pap, float* __restrict__ x,
                             const float* __restrict__ p, float* __restrict__ r, const float* __restrict__ Ap,
                             const T* __restrict__ zold, ScaleRef sc, T* __restrict__ rT, int q_rr, int q_cross) {
  const float zscale = sc.z(ws), inv_s = sc.inv_next(ws);
  double pap = ws[q_pap];
  float alpha = pap > 0.0 ? float(ws[q_rz] / pap) : 0.0f;
  double rr = 0.0, rzo = 0.0;
  FX_FOR_CELLS_N(n) {
    x[i] += alpha * p[i];
    float rv = r[i] - alpha * Ap[i];
    r[i] = rv;
    rT[i] = from_f<T>(-rv * inv_s);
    rr += double(rv) * double(rv);
    rzo += double(rv) * double(to_f(zold[i]) * zscale);
  }
  block_partial(rr, ws, q_rr);
  block_partial(rzo, ws, q_cross);
}

// partial r.(z*zscale) -> q
template <typename T>
__global__ void k_dot_rz(int64_t n, const float* __restrict__ r, const T* __restrict__ z, ScaleRef sc, double* ws,
                         int q) {
  const float zscale = sc.z(ws);
  double acc = 0.0;
  FX_FOR_CELLS_N(n) acc += double(r[i]) * double(to_f(z[i]) * zscale);
  block_partial(acc, ws, q);
}

// beta = max(0, (rz_new - rz_cross) / rz_old); p = z*zscale + beta p
template <typename T>
__global__ void k_pcg_direction(int64_t n, const double* __restrict__ ws, int q_new, int q_cross, int q_old,
                                const T* __restrict__ z, ScaleRef sc, float* __restrict__ p, int first) {
  const float zscale = sc.z(ws);
  float beta = 0.0f;
  if (!first) {
    double den = ws[q_old];
    beta = den != 0.0 ? float((ws[q_new] - ws[q_cross]) / den) : 0.0f;
    beta = fmaxf(beta, 0.0f);
  }
  FX_FOR_CELLS_N(n) p[i] = first ? to_f(z[i]) * zscale : __builtin_fmaf(beta, p[i], to_f(z[i]) * zscale);
}

// ---------------------------------------------------------------------------
// quad (four cells per lane) versions for nx % 4 == 0 and 16-byte aligned arrays
// ---------------------------------------------------------------------------
#define FX_FOR_QUADS(nq) \
  for (int q_ = blockIdx.x * blockDim.x + threadIdx.x; q_ < (nq); q_ += gridDim.x * blockDim.x)

template <int OP, typename T>
__global__ void k_ir_update_q(LevelDev L, const float* __restrict__ x, const T* __restrict__ e, float s_e,
                              const float* __restrict__ b, float* __restrict__ xo, T* __restrict__ rT, float inv_s,
                              double* ws, int q, int nq) {
  double rr = 0.0;
  FX_FOR_QUADS(nq) {
    const QuadPos at = quad_pos(L, q_);
    const QuadStencil st = quad_stencil<OP>(L, at.x, at.y, at.z, at.i, ShiftLoad<T>{x, e, s_e}, true);
    const F4 bv = ld4(b, at.i);
    F4 r, xv, rs;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool active = st.f[k] & FLAG_FLUID;
      r.v[k] = active ? bv.v[k] - st.s[k] : 0.0f;
      xv.v[k] = active ? st.c[k] : 0.0f;
      rs.v[k] = r.v[k] * inv_s;
      rr += double(r.v[k]) * double(r.v[k]);
    }
    if (xo) st4(xo, at.i, xv);
    st4(rT, at.i, rs);
  }
  block_partial(rr, ws, q);
}

template <int OP, typename T>
__global__ void k_pcg_init_q(LevelDev L, const float* __restrict__ x, const float* __restrict__ b,
                             float* __restrict__ r, T* __restrict__ rT, float inv_s, double* ws, int q, int nq) {
  double rr = 0.0;
  FX_FOR_QUADS(nq) {
    const QuadPos at = quad_pos(L, q_);
    const QuadStencil st = quad_stencil<OP>(L, at.x, at.y, at.z, at.i, PLoad<float>{x}, true);
    const F4 bv = ld4(b, at.i);
    F4 rv, rs;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      rv.v[k] = (st.f[k] & FLAG_FLUID) ? st.s[k] - bv.v[k] : 0.0f;
      rs.v[k] = -rv.v[k] * inv_s;
      rr += double(rv.v[k]) * double(rv.v[k]);
    }
    st4(r, at.i, rv);
    st4(rT, at.i, rs);
  }
  block_partial(rr, ws, q);
}

template <int OP>
__global__ void k_pcg_apply_q(LevelDev L, const float* __restrict__ p, float* __restrict__ Ap, double* ws, int q,
                              int nq) {
  double acc = 0.0;
  FX_FOR_QUADS(nq) {
    const QuadPos at = quad_pos(L, q_);
    const QuadStencil st = quad_stencil<OP>(L, at.x, at.y, at.z, at.i, PLoad<float>{p}, true);
    F4 a;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a.v[k] = (st.f[k] & FLAG_FLUID) ? -st.s[k] : 0.0f;
      acc += double(st.c[k]) * double(a.v[k]);
    }
    st4(Ap, at.i, a);
  }
  block_partial(acc, ws, q);
}

template <typename T>
__global__ void k_pcg_update_q(int nq, double* ws, int q_rz, int q_pap, float* __restrict__ x,
                               const float* __restrict__ p, float* __restrict__ r, const float* __restrict__ Ap,
                               const T* __restrict__ zold, ScaleRef sc, T* __restrict__ rT, int q_rr, int q_cross) {
  const float zscale = sc.z(ws), inv_s = sc.inv_next(ws);
  double pap = ws[q_pap];
  float alpha = pap > 0.0 ? float(ws[q_rz] / pap) : 0.0f;
  double rr = 0.0, rzo = 0.0;
  FX_FOR_QUADS(nq) {
    const int64_t i = int64_t(q_) * 4;
    F4 xv = ld4(x, i), pv = ld4(p, i), rv = ld4(r, i), av = ld4(Ap, i), zv = ld4(zold, i), rs;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      xv.v[k] += alpha * pv.v[k];
      rv.v[k] = rv.v[k] - alpha * av.v[k];
      rs.v[k] = -rv.v[k] * inv_s;
      rr += double(rv.v[k]) * double(rv.v[k]);
      rzo += double(rv.v[k]) * double(zv.v[k] * zscale);
    }
    st4(x, i, xv);
    st4(r, i, rv);
    st4(rT, i, rs);
  }
  block_partial(rr, ws, q_rr);
  block_partial(rzo, ws, q_cross);
}

template <typename T>
__global__ void k_dot_rz_q(int nq, const float* __restrict__ r, const T* __restrict__ z, ScaleRef sc, double* ws,
                           int q) {
  const float zscale = sc.z(ws);
  double acc = 0.0;
  FX_FOR_QUADS(nq) {
    const int64_t i = int64_t(q_) * 4;
    const F4 rv = ld4(r, i), zv = ld4(z, i);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc += double(rv.v[k]) * double(zv.v[k] * zscale);
  }
  block_partial(acc, ws, q);
}

template <typename T>
__global__ void k_pcg_direction_q(int nq, const double* __restrict__ ws, int q_new, int q_cross, int q_old,
                                  const T* __restrict__ z, ScaleRef sc, float* __restrict__ p, int first) {
  const float zscale = sc.z(ws);
  float beta = 0.0f;
  if (!first) {
    double den = ws[q_old];
    beta = den != 0.0 ? float((ws[q_new] - ws[q_cross]) / den) : 0.0f;
    beta = fmaxf(beta, 0.0f);
  }
  FX_FOR_QUADS(nq) {
    const int64_t i = int64_t(q_) * 4;
    const F4 zv = ld4(z, i);
    F4 pv = first ? F4{} : ld4(p, i);
#pragma unroll
    for (int k = 0; k < 4; ++k) pv.v[k] = first ? zv.v[k] * zscale : __builtin_fmaf(beta, pv.v[k], zv.v[k] * zscale);
    st4(p, i, pv);
  }
}

static const dim3 kRedGrid(kRedBlocks), kRedBlock(kRedThreads);

// r = b - L(x + s_e e) into rT (scaled) and xo; ws[q] = ||r||^2 (finalised)
static int launch_residual(const fx_mg_level& L0, int storage, const float* x, const void* e, float s_e,
                           const float* b, float* xo, void* rT, float inv_s, double* ws, int q, hipStream_t s) {
  LevelDev D = dev_level(L0);
  if (quad_ok(L0, x, b, rT) && quad_ok(L0, e, xo)) {
    const int nq = int(int64_t(L0.nx) * L0.ny * L0.nz / 4);
    FX_DISPATCH_T(storage, {
      if (L0.op == FX_OP_FLAGS)
        hipLaunchKernelGGL((k_ir_update_q<FX_OP_FLAGS, TV>), kRedGrid, kRedBlock, 0, s, D, x, (const TV*)e, s_e, b,
                           xo, (TV*)rT, inv_s, ws, q, nq);
      else
        hipLaunchKernelGGL((k_ir_update_q<FX_OP_BOX, TV>), kRedGrid, kRedBlock, 0, s, D, x, (const TV*)e, s_e, b,
                           xo, (TV*)rT, inv_s, ws, q, nq);
    });
  } else {
    FX_DISPATCH_T(storage, FX_DISPATCH_OP(L0.op, {
      hipLaunchKernelGGL((k_ir_update<OPV, TV>), kRedGrid, kRedBlock, 0, s, D, x, (const TV*)e, s_e, b, xo,
                         (TV*)rT, inv_s, ws, q);
    }));
  }
  FX_LAUNCHED("k_ir_update");
  ws_finalize(ws, q, 1, s);
  return FX_OK;
}

// reference norm^2 slot: the caller's, or ||b||^2 computed into slot 1
static int reference_slot(const fx_solve_params* prm, int64_t n, const float* b, double* ws, hipStream_t s,
                          int* ref_q) {
  *ref_q = prm->ref_q;
  if (*ref_q < 0) {
    *ref_q = 1;
    return norm2_q(n, b, ws, 1, s);
  }
  return FX_OK;
}

// read ws[0] (residual^2) and ws[ref] in one sync
static int read_residual(double* ws, int ref_q, double* rr, double* ref2, hipStream_t s) {
  double buf[kWsQuantities];
  if (int e = ws_read(ws, 0, ref_q + 1, buf, s)) return e;
  *rr = buf[0];
  *ref2 = buf[ref_q];
  return FX_OK;
}

// Flexible PCG on A = -L, c = -b.  The preconditioner fills `*z_slot`
// (storage T) from `rT` = -r/s: the V-cycle solves L e = rT, so z_A = -s*e.
struct PcgVectors {
  float* x;
  float* r;
  float* p;
  float* Ap;
  void* zold;
};

template <typename Precond>
int pcg_loop(const fx_mg_level& L0, int storage, void* rT, void** z_slot, const float* b, PcgVectors v,
             const fx_solve_params* prm, double* ws, fx_solve_info* info, Precond&& precond, hipStream_t s) {
  LevelDev D = dev_level(L0);
  const int64_t n = int64_t(L0.nx) * L0.ny * L0.nz;
  const double rms_div = 1.0 / sqrt(double(n));
  const bool fp16 = storage == FX_F16;
  // every vector (and both preconditioner slots) 16-byte aligned: quad kernels
  const bool quad = quad_ok(L0, v.x, v.r, v.p) && quad_ok(L0, v.Ap, v.zold, b) && quad_ok(L0, rT, *z_slot);
  const int nq = int(n / 4);
  int ref_q;
  if (int e = reference_slot(prm, n, b, ws, s, &ref_q)) return e;
  fx_solve_info out{};
  double rr = 0.0, ref2 = 0.0;
  const double tol2 = double(prm->tol) * double(prm->tol);
  int it = 0;
  bool done = false;
  bool measured = false;
  // slots: 0 ||r||^2, 2 r.z_old, 3 p.Ap, 4/5 alternating r.z
  for (int restart = 0; restart < 4 && !done && it < prm->max_iters; ++restart) {
    measured = true;
    int q_rz = 4, q_next = 5;
    // 16-bit residual scale: unit RMS (FP16 range); fp32/bf16 keep scale 1.  In
    // FP16 the scale of later iterations is kept on the device in the
    // alternating slots 12/13 (ScaleRef), so no iteration waits for a read-back.
    double scale = 1.0;
    int sc_cur = -1, sc_next = -1;
    if (fp16) {
      if (int e = resnorm_q(L0, v.x, b, ws, 0, 15, s)) return e;
      if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
      scale = fmax(sqrt(rr) * rms_div, 1e-30);
      sc_cur = 12;
      sc_next = 13;
      hipLaunchKernelGGL(k_set_ws, dim3(1), dim3(64), 0, s, ws, sc_cur, scale);
    }
    if (quad) {
      FX_DISPATCH_T(storage, {
        if (L0.op == FX_OP_FLAGS)
          hipLaunchKernelGGL((k_pcg_init_q<FX_OP_FLAGS, TV>), kRedGrid, kRedBlock, 0, s, D, v.x, b, v.r, (TV*)rT,
                             float(1.0 / scale), ws, 0, nq);
        else
          hipLaunchKernelGGL((k_pcg_init_q<FX_OP_BOX, TV>), kRedGrid, kRedBlock, 0, s, D, v.x, b, v.r, (TV*)rT,
                             float(1.0 / scale), ws, 0, nq);
      });
    } else {
      FX_DISPATCH_T(storage, FX_DISPATCH_OP(L0.op, {
        hipLaunchKernelGGL((k_pcg_init<OPV, TV>), kRedGrid, kRedBlock, 0, s, D, v.x, b, v.r, (TV*)rT,
                           float(1.0 / scale), ws, 0);
      }));
    }
    FX_LAUNCHED("k_pcg_init");
    ws_finalize(ws, 0, 1, s);
    if (it >= prm->check_from || restart > 0) {
      if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
      if (it == 0) out.residual0 = sqrt(rr);
      if (rr <= tol2 * ref2) {
        done = true;
        break;
      }
    }
    if (int e = precond(s)) return e;
    float zscale = float(-scale);
    // the current scale for the z consumers: the host value, or ws[sc_cur] (fp16)
    auto cur_scale = [&]() { return ScaleRef{zscale, float(1.0 / scale), sc_cur, sc_next, rms_div}; };
    // direction p = z*zscale + beta*p from slots (q_new, q_cross, q_old)
    int dir_first = 1, dq_new = q_rz, dq_cross = q_rz, dq_old = q_rz;
    // quantities of the last update (||r||^2, r.z_old) not needed before the
    // next r.z: finalised in the same launch as r.z
    uint32_t pending = 0;
    auto dot_rz = [&](int q) -> int {
      const ScaleRef sc = cur_scale();
      FX_DISPATCH_T(storage, {
        if (quad)
          hipLaunchKernelGGL((k_dot_rz_q<TV>), kRedGrid, kRedBlock, 0, s, nq, v.r, (const TV*)*z_slot, sc, ws, q);
        else
          hipLaunchKernelGGL((k_dot_rz<TV>), kRedGrid, kRedBlock, 0, s, n, v.r, (const TV*)*z_slot, sc, ws, q);
      });
      ws_finalize_mask(ws, pending | (1u << q), s);
      pending = 0;
      return FX_OK;
    };
    auto direction = [&]() -> int {
      const ScaleRef sc = cur_scale();
      FX_DISPATCH_T(storage, {
        if (quad)
          hipLaunchKernelGGL((k_pcg_direction_q<TV>), kRedGrid, kRedBlock, 0, s, nq, ws, dq_new, dq_cross, dq_old,
                             (const TV*)*z_slot, sc, v.p, dir_first);
        else
          hipLaunchKernelGGL((k_pcg_direction<TV>), kRedGrid, kRedBlock, 0, s, n, ws, dq_new, dq_cross, dq_old,
                             (const TV*)*z_slot, sc, v.p, dir_first);
      });
      return FX_OK;
    };
    if (int e = dot_rz(q_rz)) return e;
    if (int e = direction()) return e;
    FX_LAUNCHED("pcg start");
    while (it < prm->max_iters) {
      // the preconditioned residual becomes z_old; its buffer is reused next
      void* t = v.zold;
      v.zold = *z_slot;
      *z_slot = t;
      // fp32/bf16: scale 1 throughout; fp16: the update kernel derives the next
      // scale from the last finalised ||r||^2 and stores it in ws[sc_next]
      const ScaleRef upd{zscale, float(1.0 / scale), sc_cur, sc_next, rms_div};
      if (quad) {
        if (L0.op == FX_OP_FLAGS)
          hipLaunchKernelGGL((k_pcg_apply_q<FX_OP_FLAGS>), kRedGrid, kRedBlock, 0, s, D, v.p, v.Ap, ws, 3, nq);
        else
          hipLaunchKernelGGL((k_pcg_apply_q<FX_OP_BOX>), kRedGrid, kRedBlock, 0, s, D, v.p, v.Ap, ws, 3, nq);
      } else {
        FX_DISPATCH_OP(L0.op, {
          hipLaunchKernelGGL((k_pcg_apply<OPV>), kRedGrid, kRedBlock, 0, s, D, v.p, v.Ap, ws, 3);
        });
      }
      ws_finalize(ws, 3, 1, s);
      FX_DISPATCH_T(storage, {
        if (quad)
          hipLaunchKernelGGL((k_pcg_update_q<TV>), kRedGrid, kRedBlock, 0, s, nq, ws, q_rz, 3, v.x, v.p, v.r, v.Ap,
                             (const TV*)v.zold, upd, (TV*)rT, 0, 2);
        else
          hipLaunchKernelGGL((k_pcg_update<TV>), kRedGrid, kRedBlock, 0, s, n, ws, q_rz, 3, v.x, v.p, v.r, v.Ap,
                             (const TV*)v.zold, upd, (TV*)rT, 0, 2);
      });
      FX_LAUNCHED("pcg update");
      ++it;
      const uint32_t updated = (1u << 0) | (1u << 2);
      if (it < prm->check_from && it < prm->max_iters) {
        // read next by the direction (r.z_old) and the next update's fp16
        // scale (||r||^2), both after the next r.z
        pending = updated;
      } else {
        ws_finalize_mask(ws, updated, s);
      }
      if (it >= prm->check_from) {
        if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
        FX_REQUIRE(std::isfinite(rr), FX_ESTATE, "non-finite residual after %d PCG iterations", it);
        if (rr <= tol2 * ref2) break;
      }
      if (it >= prm->max_iters) break;
      if (fp16) std::swap(sc_cur, sc_next);
      if (int e = precond(s)) return e;
      if (int e = dot_rz(q_next)) return e;
      dir_first = 0;
      dq_new = q_next;
      dq_cross = 2;
      dq_old = q_rz;
      if (int e = direction()) return e;
      FX_LAUNCHED("pcg direction");
      int tmp = q_rz;
      q_rz = q_next;
      q_next = tmp;
    }
    // true residual of the final iterate (the recurrence can drift in fp32)
    if (int e = resnorm_q(L0, v.x, b, ws, 0, 15, s)) return e;
    if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
    done = rr <= tol2 * ref2;
  }
  if (!measured) {
    // no iteration budget: report the start
    if (int e = resnorm_q(L0, v.x, b, ws, 0, 15, s)) return e;
    if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
    out.residual0 = sqrt(rr);
    done = rr <= tol2 * ref2;
  }
  out.iterations = it;
  out.residual = sqrt(rr);
  out.bnorm = sqrt(ref2);
  out.converged = done ? 1 : 0;
  *info = out;
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_pcg_solve(fx_mg_config* pcfg, float* x, const float* b, float* r, float* p, float* Ap, void* zold,
                 const fx_solve_params* prm, double* ws, fx_solve_info* info, void* stream) {
  if (int e = check_config(pcfg)) return e;
  if (int e = check_params(prm)) return e;
  FX_REQUIRE(x && b && r && p && Ap && zold && ws && info, FX_EINVAL, "pcg: null array");
  FX_REQUIRE(!op_has_edges(pcfg->level[0].op), FX_EINVAL, "level 0 must be a box or flags operator");
  hipStream_t s = S(stream);
  fx_mg_level& L0 = pcfg->level[0];
  PcgVectors v{x, r, p, Ap, zold};
  auto precond = [&](hipStream_t st) { return vcycle(pcfg, 0, 1.0f, true, st); };
  return pcg_loop(L0, pcfg->storage, L0.rhs, &L0.p, b, v, prm, ws, info, precond, s);
}

int fx_learned_pcg_solve(const fx_unet* net, fx_mg_config* mg, float* x, const float* b, float* r, float* p,
                         float* Ap, float* zold, float* z, float* rT, float* ztmp, void* mg_zold,
                         const uint8_t* solid, float post_scale, int32_t learned_iters, int32_t sweeps,
                         const fx_solve_params* prm, double* ws, fx_solve_info* info, void* stream) {
  if (int e = check_config(mg)) return e;
  if (int e = check_params(prm)) return e;
  FX_REQUIRE(net && x && b && r && p && Ap && zold && z && rT && ztmp && mg_zold && ws && info, FX_EINVAL,
             "learned pcg: null array");
  FX_REQUIRE(sweeps >= 0 && sweeps % 2 == 0, FX_EINVAL, "learned pcg: smoothing sweeps must be even");
  fx_mg_level& L0 = mg->level[0];
  FX_REQUIRE(!op_has_edges(L0.op), FX_EINVAL, "level 0 must be a box or flags operator");
  // the network runs on level 0, or on level 1 (the mean-restricted residual,
  // its output prolonged as the V-cycle prolongs corrections)
  const bool coarse = net->nx * 2 == L0.nx && net->ny * 2 == L0.ny && net->nz * 2 == L0.nz;
  FX_REQUIRE(coarse || (net->nx == L0.nx && net->ny == L0.ny && net->nz == L0.nz), FX_EINVAL,
             "network grid must equal level 0 or level 1");
  FX_REQUIRE(!coarse || (mg->levels >= 2 && mg->level[1].nx == net->nx && mg->level[1].ny == net->ny &&
                         mg->level[1].nz == net->nz), FX_EINVAL, "coarse network needs multigrid level 1");
  hipStream_t s = S(stream);
  const int64_t n0 = int64_t(L0.nx) * L0.ny * L0.nz;
  fx_solve_params lp = *prm;
  lp.max_iters = std::min(prm->max_iters, std::max(0, learned_iters));
  void* zs = z;
  PcgVectors v{x, r, p, Ap, zold};
  // z = U-Net(rT) (on level 0, or prolonged from level 1), then `sweeps`
  // damped-Jacobi sweeps on L z = rT from it (the network carries the smooth
  // part of the solution, the sweeps the rough part)
  fx_mg_config sm = *mg;
  sm.storage = FX_F32;
  auto precond = [&](hipStream_t st) {
    if (coarse) {
      // ztmp is free until the sweeps: its first and second eighths hold the
      // restricted residual and the network's coarse correction
      float* rc = ztmp;
      float* zc = ztmp + n0 / 8;
      sm.level[0].p = zs;
      sm.level[0].q = ztmp;
      sm.level[0].rhs = rT;
      if (int e = fx_mg_restrict(&sm, 0, rT, rc, st)) return e;
      if (int e = unet_forward(net, rc, solid, ws, post_scale, 1, zc, st)) return e;
      if (int e = fx_mg_prolong(&sm, 0, nullptr, zc, zs, st)) return e;
    } else {
      if (int e = unet_forward(net, rT, solid, ws, post_scale, 1, (float*)zs, st)) return e;
    }
    if (sweeps <= 0) return int(FX_OK);
    sm.level[0].p = zs;
    sm.level[0].q = ztmp;
    sm.level[0].rhs = rT;
    if (int e = smooth_level(&sm, 0, sweeps, st)) return e;
    // the sweeps leave the result in either buffer: it becomes z, the other one scratch
    if (sm.level[0].p != zs) {
      ztmp = static_cast<float*>(zs);
      zs = sm.level[0].p;
    }
    return int(FX_OK);
  };
  fx_solve_info first{};
  if (int e = pcg_loop(L0, FX_F32, rT, &zs, b, v, &lp, ws, &first, precond, s)) return e;
  *info = first;
  if (first.converged || prm->max_iters <= first.iterations) return FX_OK;
  // fallback: V-cycle preconditioned FPCG from the learned iterate
  fx_solve_params fp = *prm;
  fp.max_iters = std::max(0, prm->max_iters - first.iterations);
  // host residual reads start where the whole solve's schedule puts them
  fp.check_from = std::max(0, prm->check_from - first.iterations);
  fx_solve_info second{};
  PcgVectors v2{x, r, p, Ap, mg_zold};
  auto mg_precond = [&](hipStream_t st) { return vcycle(mg, 0, 1.0f, true, st); };
  if (int e = pcg_loop(L0, mg->storage, L0.rhs, &L0.p, b, v2, &fp, ws, &second, mg_precond, s)) return e;
  second.iterations += first.iterations;
  second.residual0 = first.residual0;
  second.fallback = 1;
  *info = second;
  return FX_OK;
}

int fx_mixed_solve(fx_mg_config* ecfg, float* x, float* x_spare, const float* b, const fx_solve_params* prm,
                   double* ws, fx_solve_info* info, float** x_out, void* stream) {
  if (int e = check_config(ecfg)) return e;
  if (int e = check_params(prm)) return e;
  FX_REQUIRE(x && x_spare && b && ws && info && x_out && x != x_spare, FX_EINVAL, "mixed solve: bad arrays");
  FX_REQUIRE(!op_has_edges(ecfg->level[0].op), FX_EINVAL, "level 0 must be a box or flags operator");
  hipStream_t s = S(stream);
  fx_mg_level& L0 = ecfg->level[0];
  const int64_t n = int64_t(L0.nx) * L0.ny * L0.nz;
  const double rms_div = 1.0 / sqrt(double(n));
  const bool fp16 = ecfg->storage == FX_F16;
  int ref_q;
  if (int e = reference_slot(prm, n, b, ws, s, &ref_q)) return e;
  const double tol2 = double(prm->tol) * double(prm->tol);
  fx_solve_info out{};
  double rr = 0.0, ref2 = 0.0;
  float* cur = x;
  float* nxt = x_spare;
  int it = 0;
  bool done = false;
  // residual of the start, written as the 16-bit right-hand side
  double scale = 1.0;
  if (fp16) {
    if (int e = resnorm_q(L0, cur, b, ws, 0, 15, s)) return e;
    if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
    scale = fmax(sqrt(rr) * rms_div, 1e-30);
  }
  if (int e = launch_residual(L0, ecfg->storage, cur, nullptr, 0.0f, b, nullptr, L0.rhs, float(1.0 / scale), ws, 0,
                              s))
    return e;
  if (prm->check_from <= 0) {
    if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
    out.residual0 = sqrt(rr);
    done = rr <= tol2 * ref2;
  }
  while (!done && it < prm->max_iters) {
    if (int e = vcycle(ecfg, 0, 1.0f, true, s)) return e;
    double next_scale = fp16 ? fmax(sqrt(rr) * rms_div * 0.25, 1e-30) : 1.0;
    if (int e = launch_residual(L0, ecfg->storage, cur, L0.p, float(scale), b, nxt, L0.rhs, float(1.0 / next_scale),
                                ws, 0, s))
      return e;
    float* t = cur;
    cur = nxt;
    nxt = t;
    ++it;
    scale = next_scale;
    if (fp16 || it >= prm->check_from) {
      if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
      FX_REQUIRE(std::isfinite(rr), FX_ESTATE, "non-finite residual after %d refinement steps", it);
      done = rr <= tol2 * ref2;
      if (fp16 && !done) {
        // keep the stored 16-bit residual near unit RMS
        double actual = sqrt(rr) * rms_div;
        double ratio = actual / scale;
        if (ratio > 64.0 || ratio < 1.0 / 64.0) {
          hipLaunchKernelGGL((k_rescale<_Float16>), dim3((n + 255) / 256), dim3(256), 0, s, n, (_Float16*)L0.rhs,
                             float(1.0 / ratio));
          FX_LAUNCHED("k_rescale");
          scale = actual;
        }
      }
    }
  }
  if (!done && it < prm->check_from) {
    if (int e = read_residual(ws, ref_q, &rr, &ref2, s)) return e;
    done = rr <= tol2 * ref2;
  }
  out.iterations = it;
  out.residual = sqrt(rr);
  out.bnorm = sqrt(ref2);
  out.converged = done ? 1 : 0;
  *info = out;
  *x_out = cur;
  return FX_OK;
}

}  // extern "C"
