// FP64 reductions over level-0 fp32 arrays and the stationary
// tolerance-driven V-cycle solve.  Reduction kernels run on a fixed 1-D grid of
// kRedBlocks blocks with grid-stride loops; each block stores one partial per
// quantity and a single finalize block folds them (no atomics).
#include "quad.hpp"

#include <mutex>

namespace fx {

// One block per quantity: block b folds the partials of the b-th set bit of
// `mask` into ws[q] (the same summation order for every quantity and launch).
__global__ void k_finalize(double* ws, uint32_t mask) {
  __shared__ double part[kRedThreads / 64];
  uint32_t m = mask;
  for (unsigned k = 0; k < blockIdx.x; ++k) m &= m - 1;
  const int q = __builtin_ctz(m);
  double v = 0.0;
  for (int b = threadIdx.x; b < kRedBlocks; b += kRedThreads) v += ws[kWsBase + kRedBlocks * q + b];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kRedThreads / 64; ++k) t += part[k];
    ws[q] = t;
  }
}

void ws_finalize_mask(double* ws, uint32_t mask, hipStream_t s) {
  if (mask) hipLaunchKernelGGL(k_finalize, dim3(__builtin_popcount(mask)), dim3(kRedThreads), 0, s, ws, mask);
}

void ws_finalize(double* ws, int q0, int nq, hipStream_t s) {
  ws_finalize_mask(ws, ((nq >= 32 ? 0u : (1u << nq)) - 1u) << q0, s);
}

// Small pinned host buffer for the scalar read-backs of the solvers.
constexpr int kMaxDevices = 64;

static double* pinned(int n) {
  static thread_local double* buf = nullptr;
  if (!buf) {
    if (hipHostMalloc(reinterpret_cast<void**>(&buf), 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
      buf = nullptr;
  }
  return n <= 64 ? buf : nullptr;
}

int ws_read(const double* ws, int q0, int nq, double* host, hipStream_t s) {
  double* p = pinned(nq);
  if (p) {
    FX_HIP(hipMemcpyAsync(p, ws + q0, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nq; ++i) host[i] = p[i];
  } else {
    FX_HIP(hipMemcpyAsync(host, ws + q0, sizeof(double) * nq, hipMemcpyDeviceToHost, s));
    FX_HIP(hipStreamSynchronize(s));
  }
  return FX_OK;
}

// ws[qr] = ||b - L x||^2 (active cells), ws[qb] = ||b||^2
template <int OP>
__global__ void k_resnorm(LevelDev L, const float* __restrict__ x, const float* __restrict__ b, double* ws, int qr,
                          int qb) {
  double rr = 0.0, bb = 0.0;
  FX_FOR_ROW_CELLS(L, X, Y, Z, i) {
    float r = residual_at<float, OP>(L, X, Y, Z, x, b, 1.0f);
    float bv = b[i];
    rr += double(r) * double(r);
    bb += double(bv) * double(bv);
  }
  block_partial(rr, ws, qr);
  block_partial(bb, ws, qb);
}

template <int OP>
__global__ void k_mean_sum(LevelDev L, const float* __restrict__ b, double* ws) {
  double s = 0.0, c = 0.0;
  FX_FOR_ROW_CELLS(L, X, Y, Z, i) {
    if (cell_active<OP>(L, X, Y, Z, i)) {
      s += b[i];
      c += 1.0;
    }
  }
  block_partial(s, ws, 2);
  block_partial(c, ws, 3);
}

template <int OP>
__global__ void k_sub_mean(LevelDev L, float* __restrict__ b, const double* __restrict__ ws) {
  double cnt = ws[3];
  float mean = cnt > 0 ? float(ws[2] / cnt) : 0.0f;
  FX_FOR_ROW_CELLS(L, X, Y, Z, i) {
    b[i] = cell_active<OP>(L, X, Y, Z, i) ? b[i] - mean : 0.0f;
  }
}

// r0 = b - L x0 (neighbour differences first); ws[6] = ||r0||^2, ws[7] = ||b||^2
template <int OP>
__global__ void k_increment_rhs(LevelDev L, const float* __restrict__ x0, const float* __restrict__ b,
                                float* __restrict__ r0, double* ws) {
  double rr = 0.0, bb = 0.0;
  FX_FOR_ROW_CELLS(L, X, Y, Z, i) {
    float r = residual_at<float, OP>(L, X, Y, Z, x0, b, 1.0f);
    r0[i] = r;
    float bv = b[i];
    rr += double(r) * double(r);
    bb += double(bv) * double(bv);
  }
  if (ws) {
    block_partial(rr, ws, 6);
    block_partial(bb, ws, 7);
  }
}

// Quad versions of the level-0 reductions (nx % 4 == 0, aligned arrays).
#define FX_FOR_QUADS(nq) \
  for (int q_ = blockIdx.x * blockDim.x + threadIdx.x; q_ < (nq); q_ += gridDim.x * blockDim.x)

template <int OP>
__global__ void k_resnorm_q(LevelDev L, const float* __restrict__ x, const float* __restrict__ b, double* ws, int qr,
                            int qb, int nq) {
  double rr = 0.0, bb = 0.0;
  FX_FOR_QUADS(nq) {
    const QuadPos at = quad_pos(L, q_);
    const QuadStencil st = quad_stencil<OP>(L, at.x, at.y, at.z, at.i, PLoad<float>{x}, true);
    const F4 bv = ld4(b, at.i);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float r = (st.f[k] & FLAG_FLUID) ? bv.v[k] - st.s[k] : 0.0f;
      rr += double(r) * double(r);
      bb += double(bv.v[k]) * double(bv.v[k]);
    }
  }
  block_partial(rr, ws, qr);
  block_partial(bb, ws, qb);
}

template <int OP>
__global__ void k_increment_rhs_q(LevelDev L, const float* __restrict__ x0, const float* __restrict__ b,
                                  float* __restrict__ r0, double* ws, int nq) {
  double rr = 0.0, bb = 0.0;
  FX_FOR_QUADS(nq) {
    const QuadPos at = quad_pos(L, q_);
    const QuadStencil st = quad_stencil<OP>(L, at.x, at.y, at.z, at.i, PLoad<float>{x0}, true);
    const F4 bv = ld4(b, at.i);
    F4 r;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      r.v[k] = (st.f[k] & FLAG_FLUID) ? bv.v[k] - st.s[k] : 0.0f;
      rr += double(r.v[k]) * double(r.v[k]);
      bb += double(bv.v[k]) * double(bv.v[k]);
    }
    st4(r0, at.i, r);
  }
  if (ws) {
    block_partial(rr, ws, 6);
    block_partial(bb, ws, 7);
  }
}

template <int OP>
__global__ void k_mean_sum_q(LevelDev L, const float* __restrict__ b, double* ws, int nq) {
  double s = 0.0, c = 0.0;
  FX_FOR_QUADS(nq) {
    const QuadPos at = quad_pos(L, q_);
    const uint32_t fw = quad_flags(L, OP, at.x, at.y, at.z, at.i);
    const F4 bv = ld4(b, at.i);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if ((fw >> (8 * k)) & FLAG_FLUID) {
        s += bv.v[k];
        c += 1.0;
      }
    }
  }
  block_partial(s, ws, 2);
  block_partial(c, ws, 3);
}

template <int OP>
__global__ void k_sub_mean_q(LevelDev L, float* __restrict__ b, const double* __restrict__ ws, int nq) {
  double cnt = ws[3];
  float mean = cnt > 0 ? float(ws[2] / cnt) : 0.0f;
  FX_FOR_QUADS(nq) {
    const QuadPos at = quad_pos(L, q_);
    const uint32_t fw = quad_flags(L, OP, at.x, at.y, at.z, at.i);
    F4 bv = ld4(b, at.i);
#pragma unroll
    for (int k = 0; k < 4; ++k) bv.v[k] = ((fw >> (8 * k)) & FLAG_FLUID) ? bv.v[k] - mean : 0.0f;
    st4(b, at.i, bv);
  }
}

// Warm-start guard: a start with ||b - L x0|| >= ||b|| is worse than zero, so it
// is dropped (x0 = 0, r0 = b).  Every block reads the decision from ws[6], ws[7].
__global__ void k_drop_start(int64_t n, float* __restrict__ x0, const float* __restrict__ b,
                             float* __restrict__ r0, const double* ws) {
  if (ws[6] < ws[7]) return;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    x0[i] = 0.0f;
    r0[i] = b[i];
  }
}

// Runs after k_drop_start so that no block still reads ws[6].
__global__ void k_drop_start_norm(double* ws) {
  if (!(ws[6] < ws[7])) ws[6] = ws[7];
}

// ws[q] = ||a||^2
__global__ void k_norm2(int64_t n, const float* __restrict__ a, double* ws, int q) {
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    double v = a[i];
    s += v * v;
  }
  block_partial(s, ws, q);
}

static const dim3 kRedGrid(kRedBlocks), kRedBlock(kRedThreads);

int resnorm_q(const fx_mg_level& L0, const float* x, const float* b, double* ws, int qr, int qb, hipStream_t s) {
  LevelDev D = dev_level(L0);
  if (quad_ok(L0, x, b)) {
    const int nq = int(int64_t(L0.nx) * L0.ny * L0.nz / 4);
    if (L0.op == FX_OP_FLAGS)
      hipLaunchKernelGGL((k_resnorm_q<FX_OP_FLAGS>), kRedGrid, kRedBlock, 0, s, D, x, b, ws, qr, qb, nq);
    else
      hipLaunchKernelGGL((k_resnorm_q<FX_OP_BOX>), kRedGrid, kRedBlock, 0, s, D, x, b, ws, qr, qb, nq);
  } else {
    FX_DISPATCH_OP(L0.op, {
      hipLaunchKernelGGL((k_resnorm<OPV>), kRedGrid, kRedBlock, 0, s, D, x, b, ws, qr, qb);
    });
  }
  FX_LAUNCHED("k_resnorm");
  ws_finalize_mask(ws, (1u << qr) | (1u << qb), s);
  return FX_OK;
}

int resnorm(const fx_mg_level& L0, const float* x, const float* b, double* ws, double* rn, double* bn,
            hipStream_t s) {
  if (int e = resnorm_q(L0, x, b, ws, 0, 1, s)) return e;
  double h[2];
  if (int e = ws_read(ws, 0, 2, h, s)) return e;
  *rn = sqrt(h[0]);
  *bn = sqrt(h[1]);
  return FX_OK;
}

int norm2_q(int64_t n, const float* a, double* ws, int q, hipStream_t s) {
  hipLaunchKernelGGL(k_norm2, kRedGrid, kRedBlock, 0, s, n, a, ws, q);
  FX_LAUNCHED("k_norm2");
  ws_finalize(ws, q, 1, s);
  return FX_OK;
}

int project_mean_level(const fx_mg_level& L0, float* b, double* ws, hipStream_t s) {
  LevelDev D = dev_level(L0);
  if (quad_ok(L0, b)) {
    const int nq = int(int64_t(L0.nx) * L0.ny * L0.nz / 4);
    if (L0.op == FX_OP_FLAGS) {
      hipLaunchKernelGGL((k_mean_sum_q<FX_OP_FLAGS>), kRedGrid, kRedBlock, 0, s, D, b, ws, nq);
      ws_finalize(ws, 2, 2, s);
      hipLaunchKernelGGL((k_sub_mean_q<FX_OP_FLAGS>), kRedGrid, kRedBlock, 0, s, D, b, ws, nq);
    } else {
      hipLaunchKernelGGL((k_mean_sum_q<FX_OP_BOX>), kRedGrid, kRedBlock, 0, s, D, b, ws, nq);
      ws_finalize(ws, 2, 2, s);
      hipLaunchKernelGGL((k_sub_mean_q<FX_OP_BOX>), kRedGrid, kRedBlock, 0, s, D, b, ws, nq);
    }
  } else {
    FX_DISPATCH_OP(L0.op, {
      hipLaunchKernelGGL((k_mean_sum<OPV>), kRedGrid, kRedBlock, 0, s, D, b, ws);
      ws_finalize(ws, 2, 2, s);
      hipLaunchKernelGGL((k_sub_mean<OPV>), kRedGrid, kRedBlock, 0, s, D, b, ws);
    });
  }
  FX_LAUNCHED("k_project_mean");
  return FX_OK;
}

int check_params(const fx_solve_params* p) {
  FX_REQUIRE(p, FX_EINVAL, "null solve parameters");
  FX_REQUIRE(p->tol > 0 && p->max_iters >= 0, FX_EINVAL, "bad tolerance or iteration budget");
  FX_REQUIRE(p->ref_q >= -1 && p->ref_q < kWsQuantities, FX_EINVAL, "bad reference norm slot");
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_project_mean(fx_grid g, const uint8_t* flags, float* rhs, double* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(rhs && ws, FX_EINVAL, "null array");
  return project_mean_level(level0_of(g, flags), rhs, ws, S(stream));
}

// fx_project_mean without a caller workspace: the reduction workspace is this
// file-scope device array, one copy per GPU, found by its device address.
// Every call writes each workspace value it reads (all block partials and
// the finalised sums), so nothing needs clearing between calls; calls that
// overlap on one GPU share the array and race.
static __device__ __align__(256) double scratch[kWsDoubles];

static int scratch_on_current_device(double** out) {
  static std::mutex mu;
  static double* ptr[kMaxDevices] = {};
  int dev = -1;
  FX_HIP(hipGetDevice(&dev));
  FX_REQUIRE(dev >= 0 && dev < kMaxDevices, FX_EDEVICE, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    FX_REQUIRE(e == hipSuccess && p, FX_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
               hipGetErrorString(e));
    ptr[dev] = static_cast<double*>(p);
  }
  *out = ptr[dev];
  return FX_OK;
}

int fx_project_mean_noalloc(fx_grid g, const uint8_t* flags, float* rhs, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(rhs, FX_EINVAL, "null array");
  // the reductions fold any grid into kRedBlocks partials per quantity, so the
  // fixed workspace covers every grid check_grid accepts
  static_assert(kWsDoubles == FX_WS_DOUBLES, "scratch holds one reduction workspace");
  double* ws = nullptr;
  if (int e = scratch_on_current_device(&ws)) return e;
  return project_mean_level(level0_of(g, flags), rhs, ws, S(stream));
}

int fx_residual_norm(fx_grid g, const uint8_t* flags, const float* x, const float* b, double* rnorm, double* bnorm,
                     double* ws, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(x && b && ws && rnorm && bnorm, FX_EINVAL, "null array");
  return resnorm(level0_of(g, flags), x, b, ws, rnorm, bnorm, S(stream));
}

int fx_increment_rhs(fx_grid g, const uint8_t* flags, float* x0, const float* b, float* r0, double* ws,
                     void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(x0 && b && r0 && r0 != x0 && r0 != b, FX_EINVAL, "increment rhs: bad arrays");
  fx_mg_level L0 = level0_of(g, flags);
  LevelDev D = dev_level(L0);
  hipStream_t s = S(stream);
  if (quad_ok(L0, x0, b, r0)) {
    const int nq = int(int64_t(L0.nx) * L0.ny * L0.nz / 4);
    if (L0.op == FX_OP_FLAGS)
      hipLaunchKernelGGL((k_increment_rhs_q<FX_OP_FLAGS>), kRedGrid, kRedBlock, 0, s, D, x0, b, r0, ws, nq);
    else
      hipLaunchKernelGGL((k_increment_rhs_q<FX_OP_BOX>), kRedGrid, kRedBlock, 0, s, D, x0, b, r0, ws, nq);
  } else {
    FX_DISPATCH_OP(L0.op, {
      hipLaunchKernelGGL((k_increment_rhs<OPV>), kRedGrid, kRedBlock, 0, s, D, x0, b, r0, ws);
    });
  }
  FX_LAUNCHED("fx_increment_rhs");
  if (ws) {
    ws_finalize(ws, 6, 2, s);
    const int64_t n = int64_t(g.nx) * g.ny * g.nz;
    hipLaunchKernelGGL(k_drop_start, kRedGrid, kRedBlock, 0, s, n, x0, b, r0, (const double*)ws);
    hipLaunchKernelGGL(k_drop_start_norm, dim3(1), dim3(1), 0, s, ws);
    FX_LAUNCHED("fx_increment_rhs guard");
  }
  return FX_OK;
}

int fx_mg_solve(fx_mg_config* cfg, const fx_solve_params* prm, double* ws, fx_solve_info* info, void* stream) {
  if (int e = check_config(cfg)) return e;
  if (int e = check_params(prm)) return e;
  FX_REQUIRE(cfg->storage == FX_F32, FX_EINVAL, "fx_mg_solve runs on fp32 levels");
  FX_REQUIRE(ws && info, FX_EINVAL, "bad solve arguments");
  FX_REQUIRE(!op_has_edges(cfg->level[0].op), FX_EINVAL, "level 0 must be a box or flags operator");
  hipStream_t s = S(stream);
  fx_mg_level& L0 = cfg->level[0];
  const int64_t n = int64_t(L0.nx) * L0.ny * L0.nz;
  int ref_q = prm->ref_q;
  if (ref_q < 0) {
    ref_q = 1;
    if (int e = norm2_q(n, (const float*)L0.rhs, ws, ref_q, s)) return e;
  }
  fx_solve_info out{};
  double h[2] = {0.0, 0.0};
  bool done = false;
  int it = 0;
  // true residual of the current iterate against the reference norm
  auto check = [&]() -> int {
    if (int e = resnorm_q(L0, (const float*)L0.p, (const float*)L0.rhs, ws, 0, 15, s)) return e;
    double buf[kWsQuantities];
    if (int e = ws_read(ws, 0, ref_q + 1, buf, s)) return e;
    h[0] = buf[0];
    h[1] = buf[ref_q];
    FX_REQUIRE(std::isfinite(h[0]), FX_ESTATE, "non-finite residual after %d cycles", it);
    if (it == 0) out.residual0 = sqrt(h[0]);
    done = sqrt(h[0]) <= double(prm->tol) * sqrt(h[1]);
    return FX_OK;
  };
  if (prm->check_from <= 0)
    if (int e = check()) return e;
  while (!done && it < prm->max_iters) {
    if (int e = vcycle(cfg, 0, 1.0f, it == 0 && prm->x_is_zero, s)) return e;
    ++it;
    if (it >= prm->check_from)
      if (int e = check()) return e;
  }
  if (it < prm->check_from)
    if (int e = check()) return e;
  out.iterations = it;
  out.bnorm = sqrt(h[1]);
  out.residual = sqrt(h[0]);
  out.converged = done ? 1 : 0;
  *info = out;
  return FX_OK;
}

}  // extern "C"
