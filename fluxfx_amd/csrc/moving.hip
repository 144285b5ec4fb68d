// Moving rigid colliders (MovingSolidDevice, backend/moving_collision.py:7-87):
// owner raster, rigid wall velocity on solid faces, and the remap of cells
// and faces whose solidity changed.  The owner mask is a uint8 per cell
// (0 = fluid, k = collider k), which every solid-aware kernel reads as solid
// when nonzero.
#include "common.hpp"

namespace fx {

// affine velocity rows (3 x 4) of up to 8 colliders, by value
struct WallRows {
  float r[8][3][4];
};

// mask[c] = owner where the collider covers the cell centre and no earlier
// collider owns it (shaders/moving_mask.glsl:1-11)
__global__ void k_owner_raster(int nx, int ny, int nz, float hx, float hy, float hz, int kind, float r00, float r01,
                               float r02, float r03, float r10, float r11, float r12, float r13, float r20, float r21,
                               float r22, float r23, uint8_t owner, uint8_t* mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx * ny * nz) return;
  const int x = i % nx, r = i / nx, y = r % ny, z = r / ny;
  const float px = (x + 0.5f) * hx, py = (y + 0.5f) * hy, pz = (z + 0.5f) * hz;
  const float qx = r00 * px + r01 * py + r02 * pz + r03;
  const float qy = r10 * px + r11 * py + r12 * pz + r13;
  const float qz = r20 * px + r21 * py + r22 * pz + r23;
  const bool inside = kind == 0 ? (qx * qx + qy * qy + qz * qz <= 1.0f)
                                : (fabsf(qx) <= 1.0f && fabsf(qy) <= 1.0f && fabsf(qz) <= 1.0f);
  if (inside && mask[i] == 0) mask[i] = owner;
}

// wall velocity of interior faces of axis A next to a solid cell: the low
// cell's owner, else the high cell's, evaluated at the face centre
// (shaders/moving_wall.glsl:1-19); 0 elsewhere
template <int A>
__global__ void k_wall_velocity(int nx, int ny, int nz, float hx, float hy, float hz,
                                const uint8_t* __restrict__ owner, WallRows rows, float* __restrict__ out) {
  const int sx = nx + (A == 0), sy = ny + (A == 1), sz = nz + (A == 2);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sx * sy * sz) return;
  const int x = i % sx, r = i / sx, y = r % sy, z = r / sy;
  const int fa = A == 0 ? x : (A == 1 ? y : z), na = A == 0 ? nx : (A == 1 ? ny : nz);
  float value = 0.0f;
  if (fa > 0 && fa < na) {
    const int hi = owner[x + nx * (y + ny * z)];
    const int lo = owner[(x - (A == 0)) + nx * ((y - (A == 1)) + ny * (z - (A == 2)))];
    const int own = lo > 0 ? lo : hi;
    if (own > 0) {
      const float px = (x + (A == 0 ? 0.0f : 0.5f)) * hx, py = (y + (A == 1 ? 0.0f : 0.5f)) * hy,
                  pz = (z + (A == 2 ? 0.0f : 0.5f)) * hz;
      const float* row = rows.r[own - 1][A];
      value = row[0] * px + row[1] * py + row[2] * pz + row[3];
    }
  }
  out[i] = value;
}

// cells solid now or before start empty (shaders/moving_remap_scalar.glsl:1-7)
__global__ void k_remap_scalar(int n, const uint8_t* __restrict__ solid, const uint8_t* __restrict__ old_solid,
                               const float* __restrict__ in, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (solid[i] || old_solid[i]) ? 0.0f : in[i];
}

// faces blocked now take the wall velocity, faces blocked before start at
// rest, the others keep their value; domain-wall faces 0
// (shaders/moving_remap_face.glsl:1-12)
template <int A>
__global__ void k_remap_face(int nx, int ny, int nz, const uint8_t* __restrict__ solid,
                             const uint8_t* __restrict__ old_solid, const float* __restrict__ wall,
                             const float* __restrict__ in, float* __restrict__ out) {
  const int sx = nx + (A == 0), sy = ny + (A == 1), sz = nz + (A == 2);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sx * sy * sz) return;
  const int x = i % sx, r = i / sx, y = r % sy, z = r / sy;
  const int fa = A == 0 ? x : (A == 1 ? y : z), na = A == 0 ? nx : (A == 1 ? ny : nz);
  float value = 0.0f;
  if (fa > 0 && fa < na) {
    const int c = x + nx * (y + ny * z);
    const int l = (x - (A == 0)) + nx * ((y - (A == 1)) + ny * (z - (A == 2)));
    const bool blocked = solid[c] || solid[l], previous = old_solid[c] || old_solid[l];
    value = blocked ? wall[i] : (previous ? 0.0f : in[i]);
  }
  out[i] = value;
}

// ---------------------------------------------------------------------------
// Box-restricted update.  With moving colliders only, every solid cell before
// and after a move lies inside the box around the colliders' old and new
// poses, so outside it the mask is 0 before and after, the wall velocity 0,
// and neither the flags nor (farther than the clearance radius) the clearance
// change: the passes above, run over the box (the flags one cell wider, the
// clearance kClearanceRadius wider), give the whole-grid results.  Threads map
// to [x0, x1) x [y0, y1) x [z0 + blockIdx.z] of the cells (faces: the box's
// faces, the high face of the last cell included).
// ---------------------------------------------------------------------------
struct Box {
  int x0, y0, z0, x1, y1, z1;
};

__device__ __forceinline__ bool box_cell(const Box& b, int& x, int& y, int& z, int ex, int ey) {
  x = b.x0 + blockIdx.x * blockDim.x + threadIdx.x;
  y = b.y0 + blockIdx.y * blockDim.y + threadIdx.y;
  z = b.z0 + blockIdx.z;
  return x < b.x1 + ex && y < b.y1 + ey;
}

// owner = 0 inside the box
__global__ void k_box_clear(int nx, int ny, Box b, uint8_t* __restrict__ owner) {
  int x, y, z;
  if (!box_cell(b, x, y, z, 0, 0)) return;
  owner[idx3(x, y, z, nx, ny)] = 0;
}

__global__ void k_box_raster(int nx, int ny, Box b, float hx, float hy, float hz, int kind, WallRows rows, int k,
                             uint8_t* __restrict__ mask) {
  int x, y, z;
  if (!box_cell(b, x, y, z, 0, 0)) return;
  const float* r0 = rows.r[k][0];
  const float* r1 = rows.r[k][1];
  const float* r2 = rows.r[k][2];
  const float px = (x + 0.5f) * hx, py = (y + 0.5f) * hy, pz = (z + 0.5f) * hz;
  // the same expression as k_owner_raster, term for term
  const float qx = r0[0] * px + r0[1] * py + r0[2] * pz + r0[3];
  const float qy = r1[0] * px + r1[1] * py + r1[2] * pz + r1[3];
  const float qz = r2[0] * px + r2[1] * py + r2[2] * pz + r2[3];
  const bool inside = kind == 0 ? (qx * qx + qy * qy + qz * qz <= 1.0f)
                                : (fabsf(qx) <= 1.0f && fabsf(qy) <= 1.0f && fabsf(qz) <= 1.0f);
  const int64_t i = idx3(x, y, z, nx, ny);
  if (inside && mask[i] == 0) mask[i] = uint8_t(k + 1);
}

template <int A>
__global__ void k_box_wall_velocity(int nx, int ny, int nz, Box b, float hx, float hy, float hz,
                                    const uint8_t* __restrict__ owner, WallRows rows, float* __restrict__ out) {
  int x, y, z;
  if (!box_cell(b, x, y, z, A == 0, A == 1)) return;
  const int sx = nx + (A == 0), sy = ny + (A == 1);
  const int fa = A == 0 ? x : (A == 1 ? y : z), na = A == 0 ? nx : (A == 1 ? ny : nz);
  float value = 0.0f;
  if (fa > 0 && fa < na) {
    const int hi = owner[x + nx * (y + ny * z)];
    const int lo = owner[(x - (A == 0)) + nx * ((y - (A == 1)) + ny * (z - (A == 2)))];
    const int own = lo > 0 ? lo : hi;
    if (own > 0) {
      const float px = (x + (A == 0 ? 0.0f : 0.5f)) * hx, py = (y + (A == 1 ? 0.0f : 0.5f)) * hy,
                  pz = (z + (A == 2 ? 0.0f : 0.5f)) * hz;
      const float* row = rows.r[own - 1][A];
      value = row[0] * px + row[1] * py + row[2] * pz + row[3];
    }
  }
  out[x + int64_t(sx) * (y + int64_t(sy) * z)] = value;
}

// in place: the cells of the box solid now or before are emptied
__global__ void k_box_remap_scalars(int nx, int ny, Box b, const uint8_t* __restrict__ solid,
                                    const uint8_t* __restrict__ old_solid, float* f0, float* f1, float* f2) {
  int x, y, z;
  if (!box_cell(b, x, y, z, 0, 0)) return;
  const int64_t i = idx3(x, y, z, nx, ny);
  if (solid[i] || old_solid[i]) {
    f0[i] = 0.0f;
    if (f1) f1[i] = 0.0f;
    if (f2) f2[i] = 0.0f;
  }
}

// in place: k_remap_face's rule on the box's faces
template <int A>
__global__ void k_box_remap_face(int nx, int ny, int nz, Box b, const uint8_t* __restrict__ solid,
                                 const uint8_t* __restrict__ old_solid, const float* __restrict__ wall,
                                 float* __restrict__ f) {
  int x, y, z;
  if (!box_cell(b, x, y, z, A == 0, A == 1)) return;
  const int sx = nx + (A == 0), sy = ny + (A == 1);
  const int fa = A == 0 ? x : (A == 1 ? y : z), na = A == 0 ? nx : (A == 1 ? ny : nz);
  const int64_t i = x + int64_t(sx) * (y + int64_t(sy) * z);
  if (fa > 0 && fa < na) {
    const int c = x + nx * (y + ny * z);
    const int l = (x - (A == 0)) + nx * ((y - (A == 1)) + ny * (z - (A == 2)));
    const bool blocked = solid[c] || solid[l], previous = old_solid[c] || old_solid[l];
    if (blocked) f[i] = wall[i];
    else if (previous) f[i] = 0.0f;
  } else {
    f[i] = 0.0f;
  }
}

int check_box(const fx_grid& g, const int32_t* box, Box& b) {
  FX_REQUIRE(box, FX_EINVAL, "box: null pointer");
  b = Box{box[0], box[1], box[2], box[3], box[4], box[5]};
  FX_REQUIRE(b.x0 >= 0 && b.y0 >= 0 && b.z0 >= 0 && b.x0 <= b.x1 && b.y0 <= b.y1 && b.z0 <= b.z1 && b.x1 <= g.nx &&
                 b.y1 <= g.ny && b.z1 <= g.nz,
             FX_EINVAL, "box [%d,%d) x [%d,%d) x [%d,%d) is not inside the %d x %d x %d grid", b.x0, b.x1, b.y0, b.y1,
             b.z0, b.z1, g.nx, g.ny, g.nz);
  return FX_OK;
}

inline dim3 box_grid(const Box& b, int ex, int ey, int ez) {
  return dim3((b.x1 - b.x0 + ex + 63) / 64, (b.y1 - b.y0 + ey + 3) / 4, b.z1 - b.z0 + ez);
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_moving_update_box(fx_grid g, const int32_t* box, int32_t count, const int32_t* kinds,
                         const float* inverse_rows, const float* wall_rows, uint8_t* owner, uint8_t* flags,
                         uint8_t* clearance, uint8_t* scratch, uint8_t* scratch2, float* bu, float* bv, float* bw,
                         void* stream) {
  if (int e = check_grid(g)) return e;
  Box b;
  if (int e = check_box(g, box, b)) return e;
  FX_REQUIRE(count >= 1 && count <= 8 && kinds && inverse_rows && wall_rows, FX_EINVAL,
             "moving update: one to eight colliders with kinds and rows");
  FX_REQUIRE(owner && flags && clearance && scratch && scratch2 && bu && bv && bw, FX_EINVAL,
             "moving update: null array");
  FX_REQUIRE(scratch != scratch2 && scratch != clearance && scratch2 != clearance, FX_EINVAL,
             "moving update: the clearance scratch arrays must be distinct");
  WallRows inv{}, wall{};
  for (int k = 0; k < count; ++k) {
    FX_REQUIRE(kinds[k] == 0 || kinds[k] == 1, FX_EINVAL, "collider kind must be 0 (sphere) or 1 (box)");
    for (int a = 0; a < 3; ++a)
      for (int j = 0; j < 4; ++j) {
        inv.r[k][a][j] = inverse_rows[12 * k + 4 * a + j];
        wall.r[k][a][j] = wall_rows[12 * k + 4 * a + j];
      }
  }
  if (b.x0 == b.x1 || b.y0 == b.y1 || b.z0 == b.z1) return FX_OK;
  hipStream_t s = S(stream);
  const dim3 block(64, 4, 1);
  const float hx = g.ex / g.nx, hy = g.ey / g.ny, hz = g.ez / g.nz;
  hipLaunchKernelGGL(k_box_clear, box_grid(b, 0, 0, 0), block, 0, s, g.nx, g.ny, b, owner);
  for (int k = 0; k < count; ++k)
    hipLaunchKernelGGL(k_box_raster, box_grid(b, 0, 0, 0), block, 0, s, g.nx, g.ny, b, hx, hy, hz, kinds[k], inv, k,
                       owner);
  FX_LAUNCHED("fx_moving_update_box raster");
  const int wide[6] = {max(b.x0 - 1, 0), max(b.y0 - 1, 0), max(b.z0 - 1, 0),
                       min(b.x1 + 1, g.nx), min(b.y1 + 1, g.ny), min(b.z1 + 1, g.nz)};
  if (int e = launch_flags_box(g, wide, owner, flags, s)) return e;
  const int r = kClearanceRadius;
  const int far[6] = {max(b.x0 - r, 0), max(b.y0 - r, 0), max(b.z0 - r, 0),
                      min(b.x1 + r, g.nx), min(b.y1 + r, g.ny), min(b.z1 + r, g.nz)};
  if (int e = launch_clearance_box(g, far, owner, clearance, scratch, scratch2, s)) return e;
  hipLaunchKernelGGL(k_box_wall_velocity<0>, box_grid(b, 1, 0, 0), block, 0, s, g.nx, g.ny, g.nz, b, hx, hy, hz, owner,
                     wall, bu);
  hipLaunchKernelGGL(k_box_wall_velocity<1>, box_grid(b, 0, 1, 0), block, 0, s, g.nx, g.ny, g.nz, b, hx, hy, hz, owner,
                     wall, bv);
  // W faces: the z range of the box's faces is z0 .. z1 (the high face of the last cell)
  Box bz = b;
  bz.z1 = b.z1 + 1;
  hipLaunchKernelGGL(k_box_wall_velocity<2>, box_grid(bz, 0, 0, 0), block, 0, s, g.nx, g.ny, g.nz, bz, hx, hy, hz,
                     owner, wall, bw);
  FX_LAUNCHED("fx_moving_update_box");
  return FX_OK;
}

int fx_remap_box(fx_grid g, const int32_t* box, const uint8_t* solid, const uint8_t* old_solid, const float* bu,
                 const float* bv, const float* bw, float* u, float* v, float* w, float* s0, float* s1, float* s2,
                 void* stream) {
  if (int e = check_grid(g)) return e;
  Box b;
  if (int e = check_box(g, box, b)) return e;
  FX_REQUIRE(solid && old_solid && bu && bv && bw && u && v && w && s0, FX_EINVAL, "remap box: null array");
  if (b.x0 == b.x1 || b.y0 == b.y1 || b.z0 == b.z1) return FX_OK;
  hipStream_t s = S(stream);
  const dim3 block(64, 4, 1);
  hipLaunchKernelGGL(k_box_remap_scalars, box_grid(b, 0, 0, 0), block, 0, s, g.nx, g.ny, b, solid, old_solid, s0, s1,
                     s2);
  hipLaunchKernelGGL(k_box_remap_face<0>, box_grid(b, 1, 0, 0), block, 0, s, g.nx, g.ny, g.nz, b, solid, old_solid,
                     bu, u);
  hipLaunchKernelGGL(k_box_remap_face<1>, box_grid(b, 0, 1, 0), block, 0, s, g.nx, g.ny, g.nz, b, solid, old_solid,
                     bv, v);
  Box bz = b;
  bz.z1 = b.z1 + 1;
  hipLaunchKernelGGL(k_box_remap_face<2>, box_grid(bz, 0, 0, 0), block, 0, s, g.nx, g.ny, g.nz, bz, solid, old_solid,
                     bw, w);
  FX_LAUNCHED("fx_remap_box");
  return FX_OK;
}

int fx_owner_raster(fx_grid g, int32_t kind, const float* r, int32_t owner, uint8_t* mask, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(kind == 0 || kind == 1, FX_EINVAL, "collider kind must be 0 (sphere) or 1 (box)");
  FX_REQUIRE(owner >= 1 && owner <= 8, FX_EINVAL, "owner must be in [1, 8]");
  FX_REQUIRE(r && mask, FX_EINVAL, "null pointer");
  const int n = g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_owner_raster, dim3((n + 255) / 256), dim3(256), 0, S(stream), g.nx, g.ny, g.nz, g.ex / g.nx,
                     g.ey / g.ny, g.ez / g.nz, kind, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10],
                     r[11], uint8_t(owner), mask);
  FX_LAUNCHED("fx_owner_raster");
  return FX_OK;
}

int fx_wall_velocity(fx_grid g, const uint8_t* owner, int32_t count, const float* rows, float* bu, float* bv,
                     float* bw, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(owner && bu && bv && bw && (count == 0 || rows), FX_EINVAL, "wall velocity: null pointer");
  FX_REQUIRE(count >= 0 && count <= 8, FX_EINVAL, "wall velocity: one to eight colliders");
  WallRows w{};
  for (int k = 0; k < count; ++k)
    for (int a = 0; a < 3; ++a)
      for (int j = 0; j < 4; ++j) w.r[k][a][j] = rows[12 * k + 4 * a + j];
  const float hx = g.ex / g.nx, hy = g.ey / g.ny, hz = g.ez / g.nz;
  const int nu = (g.nx + 1) * g.ny * g.nz, nv = g.nx * (g.ny + 1) * g.nz, nw = g.nx * g.ny * (g.nz + 1);
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_wall_velocity<0>, dim3((nu + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz, hx, hy, hz,
                     owner, w, bu);
  hipLaunchKernelGGL(k_wall_velocity<1>, dim3((nv + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz, hx, hy, hz,
                     owner, w, bv);
  hipLaunchKernelGGL(k_wall_velocity<2>, dim3((nw + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz, hx, hy, hz,
                     owner, w, bw);
  FX_LAUNCHED("fx_wall_velocity");
  return FX_OK;
}

int fx_remap_scalar(fx_grid g, const uint8_t* solid, const uint8_t* old_solid, const float* in, float* out,
                    void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(solid && old_solid && in && out && in != out, FX_EINVAL, "remap: needs distinct input and output");
  const int n = g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_remap_scalar, dim3((n + 255) / 256), dim3(256), 0, S(stream), n, solid, old_solid, in, out);
  FX_LAUNCHED("fx_remap_scalar");
  return FX_OK;
}

int fx_remap_faces(fx_grid g, const uint8_t* solid, const uint8_t* old_solid, const float* bu, const float* bv,
                   const float* bw, const float* u, const float* v, const float* w, float* uo, float* vo, float* wo,
                   void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(solid && old_solid && bu && bv && bw && u && v && w && uo && vo && wo, FX_EINVAL, "remap: null array");
  FX_REQUIRE(uo != u && vo != v && wo != w, FX_EINVAL, "Advection input/output textures must be distinct");
  const int nu = (g.nx + 1) * g.ny * g.nz, nv = g.nx * (g.ny + 1) * g.nz, nw = g.nx * g.ny * (g.nz + 1);
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_remap_face<0>, dim3((nu + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz, solid, old_solid,
                     bu, u, uo);
  hipLaunchKernelGGL(k_remap_face<1>, dim3((nv + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz, solid, old_solid,
                     bv, v, vo);
  hipLaunchKernelGGL(k_remap_face<2>, dim3((nw + 255) / 256), dim3(256), 0, s, g.nx, g.ny, g.nz, solid, old_solid,
                     bw, w, wo);
  FX_LAUNCHED("fx_remap_faces");
  return FX_OK;
}

}  // extern "C"
