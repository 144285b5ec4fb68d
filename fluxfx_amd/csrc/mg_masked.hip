// Rediscretised masked multigrid operators rebuilt on the device, for solid
// masks that change between steps (moving colliders).
//
// The open fraction of a coarse face at level l is the mean of the open
// fractions of the four level-(l-1) faces it covers, so it is an exact
// multiple of 4^-l: with K the integer numerator, a coarse face's K is the sum
// of the four fine K values (a level-0 face has K = 1 when both of its cells
// are fluid).  The +a face of coarse cell C covers the +a faces of the four
// children of C on its high side along a.  A coarse cell is active when any
// child is.  Level l stores K in packed words (FX_OP_PACKED: 10 bits per axis,
// activity in bit 30) while 4^l <= 1023, else float4 conductances
// (FX_OP_EDGES: K 4^-l (n_a / extent_a)^2 per axis, evaluated in FP64 and
// rounded once, and 1/0 activity).  The same integers and roundings as the
// host construction of the static masks, so both give the same bits.
#include "mg_internal.hpp"

namespace fx {

namespace {

struct Unit {
  double w[3];  // (n_a / extent_a)^2 of the level
  double inv[3];  // 4^level / w: K from a stored conductance
  double frac;    // 4^-level
};

// K of the +x, +y, +z faces and the activity of fine cell i
template <int FOP>
__device__ __forceinline__ void fine_k(const void* op, int64_t i, const Unit& u, uint32_t& kx, uint32_t& ky,
                                       uint32_t& kz, bool& active) {
  if (FOP == FX_OP_FLAGS) {
    const uint8_t f = reinterpret_cast<const uint8_t*>(op)[i];
    kx = (f & FLAG_XHI) ? 1u : 0u;
    ky = (f & FLAG_YHI) ? 1u : 0u;
    kz = (f & FLAG_ZHI) ? 1u : 0u;
    active = f & FLAG_FLUID;
  } else if (FOP == FX_OP_PACKED) {
    const uint32_t w = reinterpret_cast<const uint32_t*>(op)[i];
    kx = w & 1023u;
    ky = (w >> 10) & 1023u;
    kz = (w >> 20) & 1023u;
    active = (w >> 30) & 1u;
  } else {
    const float4 e = reinterpret_cast<const float4*>(op)[i];
    kx = uint32_t(rint(double(e.x) * u.inv[0]));
    ky = uint32_t(rint(double(e.y) * u.inv[1]));
    kz = uint32_t(rint(double(e.z) * u.inv[2]));
    active = e.w > 0.0f;
  }
}

template <int FOP, int COP>
__global__ void k_coarsen_masked(int cx, int cy, int cz, const void* __restrict__ fine, Unit fu,
                                 void* __restrict__ coarse, Unit cu, uint8_t* __restrict__ coarse_solid) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= cx || y >= cy) return;
  const int fx = 2 * cx, fy = 2 * cy;
  uint32_t K[3] = {0u, 0u, 0u};
  bool active = false;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
    uint32_t kx, ky, kz;
    bool a;
    fine_k<FOP>(fine, idx3(2 * x + dx, 2 * y + dy, 2 * z + dz, fx, fy), fu, kx, ky, kz, a);
    if (dx) K[0] += kx;
    if (dy) K[1] += ky;
    if (dz) K[2] += kz;
    active = active || a;
  }
  const int64_t i = idx3(x, y, z, cx, cy);
  if (COP == FX_OP_PACKED) {
    reinterpret_cast<uint32_t*>(coarse)[i] = K[0] | (K[1] << 10) | (K[2] << 20) | (active ? 1u << 30 : 0u);
  } else {
    reinterpret_cast<float4*>(coarse)[i] =
        make_float4(float(double(K[0]) * cu.frac * cu.w[0]), float(double(K[1]) * cu.frac * cu.w[1]),
                    float(double(K[2]) * cu.frac * cu.w[2]), active ? 1.0f : 0.0f);
  }
  if (coarse_solid) coarse_solid[i] = active ? 0 : 1;
}

// activity bytes: bit 0 = active, bit 1 = the cell and its 26 neighbours
// (indices clamped to the level) are all active
template <int OP>
__global__ void k_level_activity(int nx, int ny, int nz, const void* __restrict__ op, uint8_t* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y, z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  auto act = [&](int ix, int iy, int iz) {
    const int64_t j = idx3(ix, iy, iz, nx, ny);
    if (OP == FX_OP_PACKED) return ((reinterpret_cast<const uint32_t*>(op)[j] >> 30) & 1u) != 0;
    return reinterpret_cast<const float4*>(op)[j].w > 0.0f;
  };
  bool all = true;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx)
        all = all && act(min(max(x + dx, 0), nx - 1), min(max(y + dy, 0), ny - 1), min(max(z + dz, 0), nz - 1));
  out[idx3(x, y, z, nx, ny)] = uint8_t((act(x, y, z) ? 1 : 0) | (all ? 2 : 0));
}

Unit level_unit(const fx_mg_level& L, int level, double ex, double ey, double ez) {
  Unit u;
  const double e[3] = {ex, ey, ez};
  const int n[3] = {L.nx, L.ny, L.nz};
  double scale = 1.0;
  for (int k = 0; k < level; ++k) scale *= 4.0;
  for (int a = 0; a < 3; ++a) {
    const double r = double(n[a]) / e[a];
    u.w[a] = r * r;
    u.inv[a] = scale / u.w[a];
  }
  u.frac = 1.0 / scale;
  return u;
}

}  // namespace

}  // namespace fx

using namespace fx;

extern "C" {

int fx_mg_rebuild_masked(const fx_mg_config* cfg, double ex, double ey, double ez, uint8_t* coarse_solid,
                         void* stream) {
  FX_REQUIRE(cfg && cfg->levels >= 2 && cfg->levels <= 16, FX_EINVAL, "rebuild: needs a config of 2 to 16 levels");
  FX_REQUIRE(ex > 0 && ey > 0 && ez > 0 && std::isfinite(ex) && std::isfinite(ey) && std::isfinite(ez), FX_EINVAL,
             "rebuild: extent must be positive and finite");
  const fx_mg_level& L0 = cfg->level[0];
  FX_REQUIRE(L0.op == FX_OP_FLAGS && L0.op_data, FX_EINVAL, "rebuild: level 0 must carry open-face flags");
  for (int l = 1; l < cfg->levels; ++l) {
    const fx_mg_level& F = cfg->level[l - 1];
    const fx_mg_level& L = cfg->level[l];
    FX_REQUIRE(F.nx == 2 * L.nx && F.ny == 2 * L.ny && F.nz == 2 * L.nz && L.nx >= 1 && L.ny >= 1 && L.nz >= 1,
               FX_EINVAL, "rebuild: level %d is not an exact factor-two coarsening", l);
    FX_REQUIRE(L.op_data && L.active, FX_EINVAL, "rebuild: level %d needs operator and activity arrays", l);
    FX_REQUIRE(L.op == FX_OP_EDGES || (L.op == FX_OP_PACKED && l <= 4), FX_EINVAL,
               "rebuild: level %d must be EDGES, or PACKED while 4^level <= 1023", l);
  }
  hipStream_t s = S(stream);
  const dim3 block(64, 4, 1);
  for (int l = 1; l < cfg->levels; ++l) {
    const fx_mg_level& F = cfg->level[l - 1];
    const fx_mg_level& L = cfg->level[l];
    const Unit fu = level_unit(F, l - 1, ex, ey, ez), cu = level_unit(L, l, ex, ey, ez);
    const dim3 grid = grid3(L.nx, L.ny, L.nz, block);
    void* out = const_cast<void*>(L.op_data);
    uint8_t* solid1 = l == 1 ? coarse_solid : nullptr;
#define FX_COARSEN(FOPV, COPV)                                                                                  \
  hipLaunchKernelGGL((k_coarsen_masked<FOPV, COPV>), grid, block, 0, s, L.nx, L.ny, L.nz, F.op_data, fu, out, cu, \
                     solid1)
    if (F.op == FX_OP_FLAGS) {
      if (L.op == FX_OP_PACKED) FX_COARSEN(FX_OP_FLAGS, FX_OP_PACKED);
      else FX_COARSEN(FX_OP_FLAGS, FX_OP_EDGES);
    } else if (F.op == FX_OP_PACKED) {
      if (L.op == FX_OP_PACKED) FX_COARSEN(FX_OP_PACKED, FX_OP_PACKED);
      else FX_COARSEN(FX_OP_PACKED, FX_OP_EDGES);
    } else {
      FX_COARSEN(FX_OP_EDGES, FX_OP_EDGES);
    }
#undef FX_COARSEN
    FX_LAUNCHED("k_coarsen_masked");
    uint8_t* act = const_cast<uint8_t*>(L.active);
    if (L.op == FX_OP_PACKED)
      hipLaunchKernelGGL((k_level_activity<FX_OP_PACKED>), grid, block, 0, s, L.nx, L.ny, L.nz, out, act);
    else
      hipLaunchKernelGGL((k_level_activity<FX_OP_EDGES>), grid, block, 0, s, L.nx, L.ny, L.nz, out, act);
    FX_LAUNCHED("k_level_activity");
  }
  return FX_OK;
}

}  // extern "C"
