// Vorticity confinement (DetailPasses.confine, backend/detail.py:39-49):
// cell-centred vorticity and its magnitude, then a bounded confinement force
// added to each face component.  One thread per cell / face, x fastest; every
// neighbour read is a clamped index, as the shaders' texelFetch of clamped
// coordinates.
#include "common.hpp"

namespace fx {

struct FaceSet {
  const float* __restrict__ u;
  const float* __restrict__ v;
  const float* __restrict__ w;
  int nx, ny, nz;
};

__device__ __forceinline__ int clampi(int a, int n) { return min(max(a, 0), n - 1); }

// centeredVelocity (shaders/curl.glsl:1-7)
__device__ __forceinline__ float3 centered_velocity(const FaceSet& F, int x, int y, int z) {
  x = clampi(x, F.nx);
  y = clampi(y, F.ny);
  z = clampi(z, F.nz);
  const int iu = x + (F.nx + 1) * (y + F.ny * z), iv = x + F.nx * (y + (F.ny + 1) * z),
            iw = x + F.nx * (y + F.ny * z);
  return make_float3(0.5f * (F.u[iu] + F.u[iu + 1]), 0.5f * (F.v[iv] + F.v[iv + F.nx]),
                     0.5f * (F.w[iw] + F.w[iw + F.nx * F.ny]));
}

// (omega, |omega|) per cell (shaders/curl.glsl:8-18)
__global__ void k_curl(FaceSet F, float ix, float iy, float iz, float4* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = i % F.nx, r = i / F.nx, y = r % F.ny, z = r / F.ny;
  const float3 xp = centered_velocity(F, x + 1, y, z), xm = centered_velocity(F, x - 1, y, z);
  const float3 yp = centered_velocity(F, x, y + 1, z), ym = centered_velocity(F, x, y - 1, z);
  const float3 zp = centered_velocity(F, x, y, z + 1), zm = centered_velocity(F, x, y, z - 1);
  const float3 dx = make_float3((xp.x - xm.x) * 0.5f * ix, (xp.y - xm.y) * 0.5f * ix, (xp.z - xm.z) * 0.5f * ix);
  const float3 dy = make_float3((yp.x - ym.x) * 0.5f * iy, (yp.y - ym.y) * 0.5f * iy, (yp.z - ym.z) * 0.5f * iy);
  const float3 dz = make_float3((zp.x - zm.x) * 0.5f * iz, (zp.y - zm.y) * 0.5f * iz, (zp.z - zm.z) * 0.5f * iz);
  const float ox = dy.z - dz.y, oy = dz.x - dx.z, oz = dx.y - dy.x;
  out[i] = make_float4(ox, oy, oz, sqrtf(ox * ox + oy * oy + oz * oz));
}

// forceAt (shaders/confinement.glsl:1-15) at cell (x, y, z), clamped reads
__device__ __forceinline__ float3 confinement_force(const float4* __restrict__ cu, int nx, int ny, int nz, float ix,
                                                    float iy, float iz, float strength, float limit, int x, int y,
                                                    int z) {
  auto at = [&](int a, int b, int c) { return cu[clampi(a, nx) + nx * (clampi(b, ny) + ny * clampi(c, nz))]; };
  const float gx = 0.5f * ix * (at(x + 1, y, z).w - at(x - 1, y, z).w);
  const float gy = 0.5f * iy * (at(x, y + 1, z).w - at(x, y - 1, z).w);
  const float gz = 0.5f * iz * (at(x, y, z + 1).w - at(x, y, z - 1).w);
  const float gm = sqrtf(gx * gx + gy * gy + gz * gz);
  if (gm < 1e-8f) return make_float3(0.0f, 0.0f, 0.0f);
  const float h = 1.0f / fmaxf(ix, fmaxf(iy, iz));
  const float4 o = at(x, y, z);
  const float nx_ = gx / gm, ny_ = gy / gm, nz_ = gz / gm;
  const float s = strength * h;
  float3 f = make_float3(s * (ny_ * o.z - nz_ * o.y), s * (nz_ * o.x - nx_ * o.z), s * (nx_ * o.y - ny_ * o.x));
  const float fm = sqrtf(f.x * f.x + f.y * f.y + f.z * f.z);
  const float k = fminf(1.0f, limit / fmaxf(fm, 1e-12f));
  return make_float3(f.x * k, f.y * k, f.z * k);
}

// face value + dt * component AXIS of the mean force of cells c and c - e_AXIS
// (shaders/confinement.glsl:16-25)
template <int AXIS>
__global__ void k_confine(const float4* __restrict__ cu, int nx, int ny, int nz, float ix, float iy, float iz,
                          const float* __restrict__ in, float* __restrict__ out, float dt, float strength,
                          float limit, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sx = nx + (AXIS == 0), sy = ny + (AXIS == 1);
  const int x = i % sx, r = i / sx, y = r % sy, z = r / sy;
  const float3 a = confinement_force(cu, nx, ny, nz, ix, iy, iz, strength, limit, x, y, z);
  const float3 b = confinement_force(cu, nx, ny, nz, ix, iy, iz, strength, limit, x - (AXIS == 0),
                                     y - (AXIS == 1), z - (AXIS == 2));
  const float fa = AXIS == 0 ? 0.5f * (a.x + b.x) : (AXIS == 1 ? 0.5f * (a.y + b.y) : 0.5f * (a.z + b.z));
  out[i] = in[i] + dt * fa;
}

// Split form: the force of every cell once (stored as three component
// planes), then each face adds the mean of its two cells' stored components;
// faces on the domain planes (one cell outside) evaluate that cell's force
// from the clamped reads directly, as the one-pass kernels do, so both forms
// agree bit for bit.
__global__ void k_cell_force(const float4* __restrict__ cu, int nx, int ny, int nz, float ix, float iy, float iz,
                             float strength, float limit, float* __restrict__ force, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = i % nx, r = i / nx, y = r % ny, z = r / ny;
  const float3 f = confinement_force(cu, nx, ny, nz, ix, iy, iz, strength, limit, x, y, z);
  force[i] = f.x;
  force[n + i] = f.y;
  force[2 * n + i] = f.z;
}

template <int AXIS>
__device__ __forceinline__ float force_comp(const float3& f) {
  return AXIS == 0 ? f.x : (AXIS == 1 ? f.y : f.z);
}

template <int AXIS>
__global__ void k_confine_split(const float4* __restrict__ cu, const float* __restrict__ force, int nx, int ny,
                                int nz, float ix, float iy, float iz, const float* __restrict__ in,
                                float* __restrict__ out, float dt, float strength, float limit, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sx = nx + (AXIS == 0), sy = ny + (AXIS == 1);
  const int x = i % sx, r = i / sx, y = r % sy, z = r / sy;
  const int c = AXIS == 0 ? x : (AXIS == 1 ? y : z), size = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
  const float* plane = force + int64_t(AXIS) * nx * ny * nz;
  const int64_t cell = x + int64_t(nx) * (y + int64_t(ny) * z);  // a cell index when c < size
  const int64_t below = cell - (AXIS == 0 ? 1 : (AXIS == 1 ? nx : int64_t(nx) * ny));
  const float a = c < size ? plane[cell]
                           : force_comp<AXIS>(confinement_force(cu, nx, ny, nz, ix, iy, iz, strength, limit, x, y, z));
  const float b = c > 0 ? plane[below]
                        : force_comp<AXIS>(confinement_force(cu, nx, ny, nz, ix, iy, iz, strength, limit,
                                                             x - (AXIS == 0), y - (AXIS == 1), z - (AXIS == 2)));
  out[i] = in[i] + dt * (0.5f * (a + b));
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_curl(fx_grid g, const float* u, const float* v, const float* w, float* curl4, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(u && v && w && curl4, FX_EINVAL, "curl: null array");
  const int n = g.nx * g.ny * g.nz;
  FaceSet F{u, v, w, g.nx, g.ny, g.nz};
  hipLaunchKernelGGL(k_curl, dim3((n + 255) / 256), dim3(256), 0, S(stream), F, g.nx / g.ex, g.ny / g.ey,
                     g.nz / g.ez, reinterpret_cast<float4*>(curl4), n);
  FX_LAUNCHED("fx_curl");
  return FX_OK;
}

int fx_confine(fx_grid g, const float* u, const float* v, const float* w, const float* curl4, float* uo, float* vo,
               float* wo, float dt, float strength, float limit, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(u && v && w && curl4 && uo && vo && wo, FX_EINVAL, "confine: null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  FX_REQUIRE(std::isfinite(strength) && strength >= 0.0f && std::isfinite(limit) && limit > 0.0f, FX_EINVAL,
             "confine: strength must be >= 0 and limit > 0");
  FX_REQUIRE((reinterpret_cast<uintptr_t>(curl4) & 15u) == 0, FX_EINVAL, "confine: curl field must be 16-byte aligned");
  const float ix = g.nx / g.ex, iy = g.ny / g.ey, iz = g.nz / g.ez;
  const float4* cu = reinterpret_cast<const float4*>(curl4);
  const int nu = (g.nx + 1) * g.ny * g.nz, nv = g.nx * (g.ny + 1) * g.nz, nw = g.nx * g.ny * (g.nz + 1);
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_confine<0>, dim3((nu + 255) / 256), dim3(256), 0, s, cu, g.nx, g.ny, g.nz, ix, iy, iz, u, uo,
                     dt, strength, limit, nu);
  hipLaunchKernelGGL(k_confine<1>, dim3((nv + 255) / 256), dim3(256), 0, s, cu, g.nx, g.ny, g.nz, ix, iy, iz, v, vo,
                     dt, strength, limit, nv);
  hipLaunchKernelGGL(k_confine<2>, dim3((nw + 255) / 256), dim3(256), 0, s, cu, g.nx, g.ny, g.nz, ix, iy, iz, w, wo,
                     dt, strength, limit, nw);
  FX_LAUNCHED("fx_confine");
  return FX_OK;
}

int fx_confine_split(fx_grid g, const float* u, const float* v, const float* w, const float* curl4, float* force3,
                     float* uo, float* vo, float* wo, float dt, float strength, float limit, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(u && v && w && curl4 && force3 && uo && vo && wo, FX_EINVAL, "confine: null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  FX_REQUIRE(std::isfinite(strength) && strength >= 0.0f && std::isfinite(limit) && limit > 0.0f, FX_EINVAL,
             "confine: strength must be >= 0 and limit > 0");
  FX_REQUIRE((reinterpret_cast<uintptr_t>(curl4) & 15u) == 0, FX_EINVAL, "confine: curl field must be 16-byte aligned");
  const float ix = g.nx / g.ex, iy = g.ny / g.ey, iz = g.nz / g.ez;
  const float4* cu = reinterpret_cast<const float4*>(curl4);
  const int n = g.nx * g.ny * g.nz;
  const int nu = (g.nx + 1) * g.ny * g.nz, nv = g.nx * (g.ny + 1) * g.nz, nw = g.nx * g.ny * (g.nz + 1);
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_cell_force, dim3((n + 255) / 256), dim3(256), 0, s, cu, g.nx, g.ny, g.nz, ix, iy, iz, strength,
                     limit, force3, n);
  hipLaunchKernelGGL(k_confine_split<0>, dim3((nu + 255) / 256), dim3(256), 0, s, cu, force3, g.nx, g.ny, g.nz, ix,
                     iy, iz, u, uo, dt, strength, limit, nu);
  hipLaunchKernelGGL(k_confine_split<1>, dim3((nv + 255) / 256), dim3(256), 0, s, cu, force3, g.nx, g.ny, g.nz, ix,
                     iy, iz, v, vo, dt, strength, limit, nv);
  hipLaunchKernelGGL(k_confine_split<2>, dim3((nw + 255) / 256), dim3(256), 0, s, cu, force3, g.nx, g.ny, g.nz, ix,
                     iy, iz, w, wo, dt, strength, limit, nw);
  FX_LAUNCHED("fx_confine_split");
  return FX_OK;
}
}  // extern "C"
