// Transport kernels: face predictor, fused velocity transport + jet relaxation
// + buoyancy, multi-channel scalar predictor and fused reverse/correct/source
// kernel, combustion and max-abs reduction.  One thread per face or cell;
// trilinear gathers stay in L1/L2 (neighbouring lanes read neighbouring data).
#include <atomic>

#include "common.hpp"

// Floating-point expressions in this file are evaluated as written: a product
// feeding a sum is fused only where an explicit fma says so.  The interior
// fast path and the general path then round alike (an implicit contraction
// could fuse an expression in one path and not in the other).
#pragma clang fp contract(off)

namespace fx {

struct GridDev {
  int nx, ny, nz;
  float ex, ey, ez;  // extent (m)
  float hx, hy, hz;  // cell size (m)
  float ix, iy, iz;  // cells per metre
  int fast;          // 0: every site takes the general path (fx_transport_fast_paths)
};

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }

// trace point p + s v, one rounding per component
__device__ __forceinline__ V3 trace(V3 p, float s, V3 v) {
  return v3(__builtin_fmaf(s, v.x, p.x), __builtin_fmaf(s, v.y, p.y), __builtin_fmaf(s, v.z, p.z));
}
// lattice coordinate x / h - off of an array whose sample i sits at (i + off) h
__device__ __forceinline__ float lat(float x, float inv_h, float off) { return __builtin_fmaf(x, inv_h, -off); }

// ---------------------------------------------------------------------------
// addressing: every transport array holds fewer than 2^30 elements and every
// z-plane fewer than 2^24 (checked on the host), so an element is a kernel-
// argument base plus a 32-bit byte offset (the loads keep the base in scalar
// registers) and coordinate products are 24-bit multiplies.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) { return uint32_t(__umul24(a, b)); }
__device__ __forceinline__ uint32_t lin(uint32_t x, uint32_t y, uint32_t z, uint32_t sx, uint32_t sy) {
  return x + mul24(sx, y + mul24(sy, z));
}
__device__ __forceinline__ float ldf(const float* __restrict__ base, uint32_t off) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void stf(float* __restrict__ base, uint32_t off, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) = v;
}

// Thread -> element (x, y, z) of an sx x sy x sz array: a block of 256
// threads covers 16 x 16 sites of one z-plane, its four waves 16 x 4 sites
// each, so a wave's sites lie within a few cells of each other (only waves
// near a wall or a solid leave the interior fast path).
// Block order: the hardware hands workgroups to the 8 XCDs round-robin.  The
// i-th block an XCD receives is given the next block of that XCD's current
// run of kXcdRun consecutive blocks (runs dealt to the XCDs in turn), so an
// XCD works through whole runs of neighbouring planes, which its own L2 keeps
// for the trilinear gathers, while the runs of all XCDs stay spread over the
// grid (balanced work where the plume is).  i: element index.
constexpr uint32_t kSiteTile = 16;
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdRun = 512;
__device__ __forceinline__ bool thread_site(uint32_t sx, uint32_t sy, uint32_t& x, uint32_t& y, uint32_t& z,
                                            uint32_t& i) {
  const uint32_t tiles_x = (sx + kSiteTile - 1) / kSiteTile;
  const uint32_t nb = gridDim.x * gridDim.y;
  uint32_t b = blockIdx.x + gridDim.x * blockIdx.y;
  if (b < nb - nb % (kXcds * kXcdRun)) {  // the last partial round keeps its order
    const uint32_t xcd = b % kXcds, k = b / kXcds;
    b = ((k / kXcdRun) * kXcds + xcd) * kXcdRun + k % kXcdRun;
  }
  const uint32_t plane = b / gridDim.x, t = b - plane * gridDim.x;
  const uint32_t ty = t / tiles_x, tx = t - ty * tiles_x;
  x = tx * kSiteTile + (threadIdx.x % kSiteTile);
  y = ty * kSiteTile + (threadIdx.x / kSiteTile);
  z = plane;
  if (x >= sx || y >= sy) return false;
  i = x + mul24(sx, y) + mul24(sx * sy, z);
  return true;
}

// ---------------------------------------------------------------------------
// face sampling: q clamped to [0,size-1], corner index min(b+o,size-1)
// ---------------------------------------------------------------------------
// Trilinear footprint: byte offset of the low corner and byte steps to the
// high corner along x/y/z (0 where the high index is clamped onto the low one).
struct Tri {
  uint32_t o, dx, dy, dz;
  float fx, fy, fz;
};

__device__ __forceinline__ Tri tri_face(int sx, int sy, int sz, float qx, float qy, float qz) {
  qx = fminf(fmaxf(qx, 0.0f), float(sx - 1));
  qy = fminf(fmaxf(qy, 0.0f), float(sy - 1));
  qz = fminf(fmaxf(qz, 0.0f), float(sz - 1));
  const float lx = floorf(qx), ly = floorf(qy), lz = floorf(qz);
  const uint32_t bx = uint32_t(lx), by = uint32_t(ly), bz = uint32_t(lz);
  Tri t;
  t.fx = qx - lx;
  t.fy = qy - ly;
  t.fz = qz - lz;
  t.o = lin(bx, by, bz, sx, sy) << 2;
  t.dx = bx + 1u < uint32_t(sx) ? 4u : 0u;
  t.dy = by + 1u < uint32_t(sy) ? uint32_t(sx) << 2 : 0u;
  t.dz = bz + 1u < uint32_t(sz) ? mul24(uint32_t(sx), uint32_t(sy)) << 2 : 0u;
  return t;
}

// the eight donors in the order (x, y, z) = 000, 100, 010, 110, 001, 101, 011, 111
__device__ __forceinline__ void tri_donors(const float* __restrict__ F, const Tri& t, float (&d)[8]) {
  const uint32_t o1 = t.o + t.dx, o2 = t.o + t.dy, o3 = o2 + t.dx;
  d[0] = ldf(F, t.o);
  d[1] = ldf(F, o1);
  d[2] = ldf(F, o2);
  d[3] = ldf(F, o3);
  d[4] = ldf(F, t.o + t.dz);
  d[5] = ldf(F, o1 + t.dz);
  d[6] = ldf(F, o2 + t.dz);
  d[7] = ldf(F, o3 + t.dz);
}

// trilinear value as nested linear interpolations (x, then y, then z)
__device__ __forceinline__ float lerpf(float a, float b, float t) { return __builtin_fmaf(t, b - a, a); }
__device__ __forceinline__ float tri_blend(const Tri& t, const float (&d)[8]) {
  const float c00 = lerpf(d[0], d[1], t.fx), c10 = lerpf(d[2], d[3], t.fx);
  const float c01 = lerpf(d[4], d[5], t.fx), c11 = lerpf(d[6], d[7], t.fx);
  return lerpf(lerpf(c00, c10, t.fy), lerpf(c01, c11, t.fy), t.fz);
}

__device__ __forceinline__ float sample_face(const float* __restrict__ F, int sx, int sy, int sz, float qx, float qy,
                                             float qz) {
  float d[8];
  const Tri t = tri_face(sx, sy, sz, qx, qy, qz);
  tri_donors(F, t, d);
  return tri_blend(t, d);
}

// sample_face plus the min/max of its eight donors (MacCormack clamp; the
// donor set equals the clamp set whenever q lies inside the array)
__device__ __forceinline__ float sample_face_mm(const float* __restrict__ F, int sx, int sy, int sz, float qx, float qy,
                                                float qz, float& lo, float& hi) {
  float d[8];
  const Tri t = tri_face(sx, sy, sz, qx, qy, qz);
  tri_donors(F, t, d);
  lo = fminf(fminf(fminf(d[0], d[1]), fminf(d[2], d[3])), fminf(fminf(d[4], d[5]), fminf(d[6], d[7])));
  hi = fmaxf(fmaxf(fmaxf(d[0], d[1]), fmaxf(d[2], d[3])), fmaxf(fmaxf(d[4], d[5]), fmaxf(d[6], d[7])));
  return tri_blend(t, d);
}

struct Vel {
  const float* u;
  const float* v;
  const float* w;
};

// velocityAt: q = p*cellCount/extent, three staggered samples
__device__ __forceinline__ V3 velocity_at(const GridDev& g, const Vel& V, V3 p) {
  float qx = p.x * g.ix, qy = p.y * g.iy, qz = p.z * g.iz;
  return v3(sample_face(V.u, g.nx + 1, g.ny, g.nz, qx, qy - 0.5f, qz - 0.5f),
            sample_face(V.v, g.nx, g.ny + 1, g.nz, qx - 0.5f, qy, qz - 0.5f),
            sample_face(V.w, g.nx, g.ny, g.nz + 1, qx - 0.5f, qy - 0.5f, qz));
}

// velocityAt at face (x, y, z) of axis AXIS.  On the lattice the trilinear
// weights collapse: the own component is the face value itself; another
// component b sits half a cell off along AXIS and along b and on the lattice
// along the third axis, so it is the mean of four values (indices clamped as
// the sampler clamps its coordinates).
template <int AXIS>
__device__ __forceinline__ V3 velocity_at_face(const GridDev& g, const Vel& V, int x, int y, int z, uint32_t ib) {
  auto four = [&](const float* F, int b) {
    // F: component b, dims (nx + (b==0), ny + (b==1), nz + (b==2)); lattice coords of this face in F
    const int sx = g.nx + (b == 0), sy = g.ny + (b == 1), sz = g.nz + (b == 2);
    int c[3] = {x, y, z};
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int size = a == 0 ? sx : (a == 1 ? sy : sz);
      if (a == AXIS) {          // position i_a, F coordinate i_a - 0.5
        lo[a] = max(c[a] - 1, 0);
        hi[a] = min(c[a], size - 1);
      } else if (a == b) {      // position c + 0.5, F coordinate c + 0.5
        lo[a] = min(c[a], size - 1);
        hi[a] = min(c[a] + 1, size - 1);
      } else {                  // position c + 0.5, F coordinate c (exact)
        lo[a] = hi[a] = min(c[a], size - 1);
      }
    }
    // the two half-offset axes are AXIS and b (a1 < a2); the four values are
    // summed in the order lo, +a1, +a2, +a1+a2
    const int a1 = AXIS < b ? AXIS : b, a2 = AXIS < b ? b : AXIS;
    const uint32_t stride[3] = {4u, uint32_t(sx) << 2, mul24(uint32_t(sx), uint32_t(sy)) << 2};
    const uint32_t o = lin(lo[0], lo[1], lo[2], sx, sy) << 2;
    const uint32_t s1 = hi[a1] != lo[a1] ? stride[a1] : 0u, s2 = hi[a2] != lo[a2] ? stride[a2] : 0u;
    return 0.25f * (ldf(F, o) + ldf(F, o + s1) + ldf(F, o + s2) + ldf(F, o + s1 + s2));
  };
  const float u = AXIS == 0 ? ldf(V.u, ib) : four(V.u, 0);
  const float v = AXIS == 1 ? ldf(V.v, ib) : four(V.v, 1);
  const float w = AXIS == 2 ? ldf(V.w, ib) : four(V.w, 2);
  return v3(u, v, w);
}

// velocityAt at the centre of cell (x, y, z): every component is the mean of
// the two faces of the cell along its axis
__device__ __forceinline__ V3 velocity_at_cell(const GridDev& g, const Vel& V, int x, int y, int z) {
  const uint32_t iu = lin(x, y, z, g.nx + 1, g.ny) << 2;
  const uint32_t iv = lin(x, y, z, g.nx, g.ny + 1) << 2;
  const uint32_t iw = lin(x, y, z, g.nx, g.ny) << 2;
  return v3(0.5f * (ldf(V.u, iu) + ldf(V.u, iu + 4u)), 0.5f * (ldf(V.v, iv) + ldf(V.v, iv + (uint32_t(g.nx) << 2))),
            0.5f * (ldf(V.w, iw) + ldf(V.w, iw + (mul24(uint32_t(g.nx), uint32_t(g.ny)) << 2))));
}

// ---------------------------------------------------------------------------
// solids
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool solid_cell(const GridDev& g, const uint8_t* __restrict__ s, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) return true;
  return s[lin(x, y, z, g.nx, g.ny)] != 0;
}

// 3-D DDA clip of origin->target at the first solid cell.  A trace that is
// neither clamped to the domain nor blocked returns `target` itself.
__device__ V3 clip_trace(const GridDev& g, const uint8_t* __restrict__ s, V3 o, V3 t) {
  const float cx = float(g.nx), cy = float(g.ny), cz = float(g.nz);
  float ax = fminf(fmaxf(o.x * g.ix, 1e-4f), cx - 1e-4f);
  float ay = fminf(fmaxf(o.y * g.iy, 1e-4f), cy - 1e-4f);
  float az = fminf(fmaxf(o.z * g.iz, 1e-4f), cz - 1e-4f);
  float rbx = t.x * g.ix, rby = t.y * g.iy, rbz = t.z * g.iz;
  float bx = fminf(fmaxf(rbx, 1e-4f), cx - 1e-4f);
  float by = fminf(fmaxf(rby, 1e-4f), cy - 1e-4f);
  float bz = fminf(fmaxf(rbz, 1e-4f), cz - 1e-4f);
  bool clamped = bx != rbx || by != rby || bz != rbz;
  float dx = bx - ax, dy = by - ay, dz = bz - az;
  int ix = int(floorf(ax)), iy = int(floorf(ay)), iz = int(floorf(az));
  if (solid_cell(g, s, ix, iy, iz)) return o;
  int sx = dx > 0.0f ? 1 : (dx < 0.0f ? -1 : 0);
  int sy = dy > 0.0f ? 1 : (dy < 0.0f ? -1 : 0);
  int sz = dz > 0.0f ? 1 : (dz < 0.0f ? -1 : 0);
  float crx = 1e20f, cry = 1e20f, crz = 1e20f, stx = 1e20f, sty = 1e20f, stz = 1e20f;
  if (fabsf(dx) > 1e-10f) { crx = (float(ix + (sx > 0 ? 1 : 0)) - ax) / dx; stx = 1.0f / fabsf(dx); }
  if (fabsf(dy) > 1e-10f) { cry = (float(iy + (sy > 0 ? 1 : 0)) - ay) / dy; sty = 1.0f / fabsf(dy); }
  if (fabsf(dz) > 1e-10f) { crz = (float(iz + (sz > 0 ? 1 : 0)) - az) / dz; stz = 1.0f / fabsf(dz); }
  float len = sqrtf(dx * dx + dy * dy + dz * dz);
  for (int i = 0; i < 384; ++i) {
    int axis = (crx <= cry && crx <= crz) ? 0 : (cry <= crz ? 1 : 2);
    float tt = axis == 0 ? crx : (axis == 1 ? cry : crz);
    if (tt >= 1.0f) {
      if (!clamped) return t;
      return v3(bx * g.hx, by * g.hy, bz * g.hz);
    }
    if (axis == 0) ix += sx; else if (axis == 1) iy += sy; else iz += sz;
    if (solid_cell(g, s, ix, iy, iz)) {
      float f = fmaxf(0.0f, tt - 1e-4f / fmaxf(len, 1.0f));
      return v3((ax + dx * f) * g.hx, (ay + dy * f) * g.hy, (az + dz * f) * g.hz);
    }
    if (axis == 0) crx += stx; else if (axis == 1) cry += sty; else crz += stz;
  }
  return o;
}

__device__ __forceinline__ bool solid_halo(const GridDev& g, const uint8_t* __restrict__ s, V3 p) {
  int cx = int(floorf(p.x * g.ix)), cy = int(floorf(p.y * g.iy)), cz = int(floorf(p.z * g.iz));
  for (int z = -1; z <= 1; ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x)
        if (solid_cell(g, s, cx + x, cy + y, cz + z)) return true;
  return false;
}

__device__ __forceinline__ float dist3(V3 a, V3 b) {
  float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

__device__ bool safe_correction(const GridDev& g, const uint8_t* __restrict__ s, V3 p, V3 dep, V3 mid, V3 arr) {
  if (solid_halo(g, s, p) || solid_halo(g, s, dep) || solid_halo(g, s, arr)) return false;
  if (dist3(clip_trace(g, s, p, dep), dep) >= 1e-7f) return false;
  if (dist3(clip_trace(g, s, p, mid), mid) >= 1e-7f) return false;
  return dist3(clip_trace(g, s, p, arr), arr) < 1e-7f;
}

// Clearance: Chebyshev distance (cells) from a cell to the nearest solid or
// out-of-range cell; distances above 15 are stored as 255 ("at least 16").
// A 3x3x3 halo is solid-free exactly when the clearance exceeds 1.
__device__ __forceinline__ int clearance_at(const GridDev& g, const uint8_t* __restrict__ cl, V3 p) {
  int cx = int(floorf(p.x * g.ix)), cy = int(floorf(p.y * g.iy)), cz = int(floorf(p.z * g.iz));
  if (cx < 0 || cy < 0 || cz < 0 || cx >= g.nx || cy >= g.ny || cz >= g.nz) return 0;
  return cl[lin(cx, cy, cz, g.nx, g.ny)];
}

// Solid tests from the clearance c = min(clearance, 16) of one origin cell
// (c < 0: no clearance field).  A segment o -> t of Chebyshev length r cells
// with c > ceil(r) + 1 crosses no solid (its cells lie within ceil(r) of the
// origin cell), and, the distance being 1-Lipschitz, every cell within one
// cell of t is fluid and in range: the DDA clip returns t, and the halo test
// of t fails.
struct Reach {
  float c;
  __device__ __forceinline__ bool free(const GridDev& g, V3 o, V3 t) const {
    const float r = fmaxf(fmaxf(fabsf((t.x - o.x) * g.ix), fabsf((t.y - o.y) * g.iy)), fabsf((t.z - o.z) * g.iz));
    return c > ceilf(r) + 1.0f;
  }
};

// The trace origin of a face or cell site is given by its integer cell: a
// face lies on the low boundary of cell (x, y, z), and floor() of its
// floating-point position could fall a rounding step into the cell below.
__device__ __forceinline__ int clearance_cell(const GridDev& g, const uint8_t* __restrict__ cl, int cx, int cy,
                                              int cz) {
  if (cx < 0 || cy < 0 || cz < 0 || cx >= g.nx || cy >= g.ny || cz >= g.nz) return 0;
  return cl[lin(cx, cy, cz, g.nx, g.ny)];
}

__device__ __forceinline__ Reach reach_cell(const GridDev& g, const uint8_t* __restrict__ cl, int cx, int cy,
                                            int cz) {
  return Reach{cl ? float(min(clearance_cell(g, cl, cx, cy, cz), 16)) : -1.0f};
}

__device__ __forceinline__ bool halo_cell(const GridDev& g, const uint8_t* __restrict__ s,
                                          const uint8_t* __restrict__ cl, int cx, int cy, int cz) {
  if (cl) return clearance_cell(g, cl, cx, cy, cz) <= 1;
  for (int z = -1; z <= 1; ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x)
        if (solid_cell(g, s, cx + x, cy + y, cz + z)) return true;
  return false;
}

__device__ __forceinline__ V3 clip_fast(const GridDev& g, const uint8_t* __restrict__ s, const Reach& R, V3 o,
                                        V3 t) {
  return R.free(g, o, t) ? t : clip_trace(g, s, o, t);
}

__device__ __forceinline__ bool halo_fast(const GridDev& g, const uint8_t* __restrict__ s,
                                          const uint8_t* __restrict__ cl, V3 p) {
  if (cl) return clearance_at(g, cl, p) <= 1;
  return solid_halo(g, s, p);
}

// the departure/arrival conditions of safe_correction with origin p (R its reach)
__device__ bool safe_fast(const GridDev& g, const uint8_t* __restrict__ s, const uint8_t* __restrict__ cl,
                          const Reach& R, V3 p, int pcx, int pcy, int pcz, V3 dep, V3 mid, V3 arr) {
  if (R.free(g, p, dep) && R.free(g, p, mid) && R.free(g, p, arr)) return true;
  if (halo_cell(g, s, cl, pcx, pcy, pcz) || halo_fast(g, s, cl, dep) || halo_fast(g, s, cl, arr)) return false;
  if (dist3(clip_fast(g, s, R, p, dep), dep) >= 1e-7f) return false;
  if (dist3(clip_fast(g, s, R, p, mid), mid) >= 1e-7f) return false;
  return dist3(clip_fast(g, s, R, p, arr), arr) < 1e-7f;
}

// unclipped trace p -> t whose end point may be corrected (halo fluid)
__device__ __forceinline__ bool end_free(const GridDev& g, const uint8_t* __restrict__ s,
                                         const uint8_t* __restrict__ cl, const Reach& R, V3 p, V3 t) {
  if (R.free(g, p, t)) return true;
  return !halo_fast(g, s, cl, t) && dist3(clip_fast(g, s, R, p, t), t) < 1e-7f;
}

// ---------------------------------------------------------------------------
// Interior fast path.  Most sites lie at least 4 cells from every solid and
// every wall, and their traces end within 2 cells of the site (CFL): there
// every trilinear corner is in range and every solid test passes, so the
// clamps, the solid tests and the bounds tests are known.  A wave takes the
// fast path when all its lanes qualify (tested before each gather, so no lane
// reads out of range), else the general path; both evaluate the same
// operations on the same values, so they give the same bits.
// ---------------------------------------------------------------------------
// site cell (x, y, z) at least 4 cells from every solid and wall: clearance
// >= 4 (the clearance counts out-of-range cells as solid), or without solids
// the cell index in [3, n - 4] on every axis
template <bool SOLID>
__device__ __forceinline__ bool open_site(const GridDev& g, const Reach& R, int x, int y, int z) {
  if (!g.fast) return false;
  if (SOLID) return R.c >= 4.0f;
  return x >= 3 && y >= 3 && z >= 3 && x <= g.nx - 4 && y <= g.ny - 4 && z <= g.nz - 4;
}

// t within 2 cells (Chebyshev) of the site position o.  With an open site
// (i) every trilinear corner of a face or cell sample at t is in range: the
// coordinate stays at least 1 from both ends of the sampled array, and
// (ii) Reach.free(o, t) holds: ceil(r) + 1 <= 3 < 4 <= c.
__device__ __forceinline__ bool near_site(const GridDev& g, V3 o, V3 t) {
  return fmaxf(fmaxf(fabsf((t.x - o.x) * g.ix), fabsf((t.y - o.y) * g.iy)), fabsf((t.z - o.z) * g.iz)) <= 2.0f;
}

// tri_face for q at least 1 from both ends of every axis: the clamps are
// identities and the high corner is in range
__device__ __forceinline__ Tri tri_face_in(uint32_t sx, uint32_t sy, float qx, float qy, float qz) {
  const float lx = floorf(qx), ly = floorf(qy), lz = floorf(qz);
  Tri t;
  t.fx = qx - lx;
  t.fy = qy - ly;
  t.fz = qz - lz;
  t.o = lin(uint32_t(lx), uint32_t(ly), uint32_t(lz), sx, sy) << 2;
  t.dx = 4u;
  t.dy = sx << 2;
  t.dz = mul24(sx, sy) << 2;
  return t;
}

__device__ __forceinline__ float sample_face_in(const float* __restrict__ F, uint32_t sx, uint32_t sy, float qx,
                                                float qy, float qz) {
  float d[8];
  const Tri t = tri_face_in(sx, sy, qx, qy, qz);
  tri_donors(F, t, d);
  return tri_blend(t, d);
}

__device__ __forceinline__ float sample_face_mm_in(const float* __restrict__ F, uint32_t sx, uint32_t sy, float qx,
                                                   float qy, float qz, float& lo, float& hi) {
  float d[8];
  const Tri t = tri_face_in(sx, sy, qx, qy, qz);
  tri_donors(F, t, d);
  lo = fminf(fminf(fminf(d[0], d[1]), fminf(d[2], d[3])), fminf(fminf(d[4], d[5]), fminf(d[6], d[7])));
  hi = fmaxf(fmaxf(fmaxf(d[0], d[1]), fmaxf(d[2], d[3])), fmaxf(fmaxf(d[4], d[5]), fmaxf(d[6], d[7])));
  return tri_blend(t, d);
}

// velocity_at for a point within 2 cells of an open site.  The 24 donor
// loads are issued together (the scheduling barrier keeps the compiler from
// waiting on one component's donors before loading the next component's), so
// the sample costs one memory round trip.
__device__ __forceinline__ V3 velocity_at_in(const GridDev& g, const Vel& V, V3 p) {
  float qx = p.x * g.ix, qy = p.y * g.iy, qz = p.z * g.iz;
  const Tri tu = tri_face_in(g.nx + 1, g.ny, qx, qy - 0.5f, qz - 0.5f);
  const Tri tv = tri_face_in(g.nx, g.ny + 1, qx - 0.5f, qy, qz - 0.5f);
  const Tri tw = tri_face_in(g.nx, g.ny, qx - 0.5f, qy - 0.5f, qz);
  float du[8], dv[8], dw[8];
  tri_donors(V.u, tu, du);
  tri_donors(V.v, tv, dv);
  tri_donors(V.w, tw, dw);
  __builtin_amdgcn_sched_barrier(0);
  return v3(tri_blend(tu, du), tri_blend(tv, dv), tri_blend(tw, dw));
}

// velocity_at_face at an open site: every neighbour index is in range
template <int AXIS>
__device__ __forceinline__ V3 velocity_at_face_in(const GridDev& g, const Vel& V, int x, int y, int z, uint32_t ib) {
  auto four = [&](const float* F, int b) {
    const uint32_t sx = g.nx + (b == 0), sy = g.ny + (b == 1);
    int lo[3] = {x, y, z};
    lo[AXIS] -= 1;  // F coordinate i_a - 0.5 along AXIS: cells i_a - 1 and i_a
    const int a1 = AXIS < b ? AXIS : b, a2 = AXIS < b ? b : AXIS;
    const uint32_t stride[3] = {4u, sx << 2, mul24(sx, sy) << 2};
    const uint32_t o = lin(lo[0], lo[1], lo[2], sx, sy) << 2;
    const uint32_t s1 = stride[a1], s2 = stride[a2];
    return 0.25f * (ldf(F, o) + ldf(F, o + s1) + ldf(F, o + s2) + ldf(F, o + s1 + s2));
  };
  const float u = AXIS == 0 ? ldf(V.u, ib) : four(V.u, 0);
  const float v = AXIS == 1 ? ldf(V.v, ib) : four(V.v, 1);
  const float w = AXIS == 2 ? ldf(V.w, ib) : four(V.w, 2);
  return v3(u, v, w);
}

// separable L-infinity distance transform, one axis per pass:
// pass 0: d = min |k| over solid/out-of-range cells c+k*e_x (|k| <= 254)
// pass 1,2: d = min over |k| of max(|k|, d_prev[c+k*e_a]), out-of-range -> |k|
// over the cells [x0, x1) x [y0, y1) x [z0 + blockIdx.z] of the grid
__global__ void k_clearance(int nx, int ny, int nz, int axis, const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                            int radius, int x0, int y0, int z0, int x1, int y1) {
  int x = x0 + blockIdx.x * blockDim.x + threadIdx.x, y = y0 + blockIdx.y * blockDim.y + threadIdx.y,
      z = z0 + blockIdx.z;
  if (x >= x1 || y >= y1) return;
  int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  int c = axis == 0 ? x : (axis == 1 ? y : z);
  int64_t stride = axis == 0 ? 1 : (axis == 1 ? int64_t(nx) : int64_t(nx) * ny);
  int64_t i = idx3(x, y, z, nx, ny);
  int best = 255;
  for (int k = -radius; k <= radius; ++k) {
    int ak = k < 0 ? -k : k;
    if (ak >= best) continue;
    int cc = c + k;
    int v;
    if (cc < 0 || cc >= n) {
      v = ak;
    } else {
      uint8_t src = in[i + k * stride];
      v = axis == 0 ? (src ? 0 : 255) : src;
      if (v < 255) v = max(v, ak);
    }
    best = min(best, v);
  }
  out[i] = uint8_t(best);
}

// ---------------------------------------------------------------------------
// emission: exact time-average of the compact sphere profile along the sweep
// ---------------------------------------------------------------------------
__device__ float emission_weight(V3 p, const float* row) {
  V3 c = v3(row[0], row[1], row[2]);
  float radius = row[3];
  V3 st = v3(row[4], row[5], row[6]);
  bool solid_profile = row[14] > 0.5f;
  // A point more than 2 radii outside the box around the swept segment lies
  // at least 2 radii from the segment: the tests below give 0 for it with a
  // margin far beyond their rounding, so it is answered here.
  const float m = 2.0f * radius;
  if (p.x < fminf(c.x, st.x) - m || p.x > fmaxf(c.x, st.x) + m || p.y < fminf(c.y, st.y) - m ||
      p.y > fmaxf(c.y, st.y) + m || p.z < fminf(c.z, st.z) - m || p.z > fmaxf(c.z, st.z) + m)
    return 0.0f;
  float sgx = c.x - st.x, sgy = c.y - st.y, sgz = c.z - st.z;
  float l2 = sgx * sgx + sgy * sgy + sgz * sgz;
  float r2 = radius * radius;
  if (l2 < 1e-12f) {
    float dx = p.x - c.x, dy = p.y - c.y, dz = p.z - c.z;
    float d2 = dx * dx + dy * dy + dz * dz;
    float w = fmaxf(1.0f - d2 / r2, 0.0f);
    return solid_profile ? (d2 < r2 ? 1.0f : 0.0f) : w * w;
  }
  float lp = sqrtf(l2);
  float axx = sgx / lp, axy = sgy / lp, axz = sgz / lp;
  float ox = p.x - st.x, oy = p.y - st.y, oz = p.z - st.z;
  float along = ox * axx + oy * axy + oz * axz;
  float px = ox - along * axx, py = oy - along * axy, pz = oz - along * axz;
  float a = 1.0f - (px * px + py * py + pz * pz) / r2;
  if (a <= 0.0f) return 0.0f;
  float hw = radius * sqrtf(a);
  float lo = fmaxf(-along, -hw), hi = fminf(lp - along, hw);
  if (hi <= lo) return 0.0f;
  if (solid_profile) return (hi - lo) / lp;
  float u = lo / hw, v = hi / hw;
  float u3 = u * u * u, v3_ = v * v * v;
  float integral = (v - u) - 2.0f / 3.0f * (v3_ - u3) + 0.2f * (v3_ * v * v - u3 * u * u);
  return fminf(fmaxf(a * a * hw / lp * integral, 0.0f), 1.0f);
}

// ---------------------------------------------------------------------------
// face transport: one thread per site (x, y, z) of the (nx + 1) x (ny + 1) x
// (nz + 1) lattice handles the U, V and W faces with that index (each where it
// exists), so a thread carries three independent gather chains and a wave of
// open sites takes the interior fast path for all three at once.
// ---------------------------------------------------------------------------
template <int AXIS>
struct FaceAxis {
  static constexpr float offx = AXIS == 0 ? 0.0f : 0.5f, offy = AXIS == 1 ? 0.0f : 0.5f,
                         offz = AXIS == 2 ? 0.0f : 0.5f;
  static __device__ __forceinline__ int sx(const GridDev& g) { return g.nx + (AXIS == 0); }
  static __device__ __forceinline__ int sy(const GridDev& g) { return g.ny + (AXIS == 1); }
  static __device__ __forceinline__ int sz(const GridDev& g) { return g.nz + (AXIS == 2); }
  static __device__ __forceinline__ bool exists(const GridDev& g, int x, int y, int z) {
    return x < sx(g) && y < sy(g) && z < sz(g);
  }
  // byte offset of face (x, y, z)
  static __device__ __forceinline__ uint32_t at(const GridDev& g, int x, int y, int z) {
    return lin(x, y, z, sx(g), sy(g)) << 2;
  }
  static __device__ __forceinline__ V3 pos(const GridDev& g, int x, int y, int z) {
    return v3((x + offx) * g.hx, (y + offy) * g.hy, (z + offz) * g.hz);
  }
  static __device__ __forceinline__ float qx(const GridDev& g, V3 t) { return lat(t.x, g.ix, offx); }
  static __device__ __forceinline__ float qy(const GridDev& g, V3 t) { return lat(t.y, g.iy, offy); }
  static __device__ __forceinline__ float qz(const GridDev& g, V3 t) { return lat(t.z, g.iz, offz); }
};

struct FaceIO {
  const float* in[3];
  float* out[3];
  float* lo[3];
  float* hi[3];
};

// face predictor (transport_face), general path of one face.
// BOUNDS: also store the donor min/max of the departure sample when that
// sample may be MacCormack-corrected (departure inside the face lattice, its
// halo fluid and the trace to it unclipped), else lo = +inf, hi = -inf
template <int AXIS, bool SOLID, bool BOUNDS>
__device__ void face_predict_general(const GridDev& g, const Vel& V, float dt, const FaceIO& F,
                                     const uint8_t* __restrict__ solid, const uint8_t* __restrict__ clear,
                                     const Reach& R, int x, int y, int z) {
  using FA = FaceAxis<AXIS>;
  const int sx = FA::sx(g), sy = FA::sy(g), sz = FA::sz(g);
  const uint32_t i = FA::at(g, x, y, z);
  const V3 p = FA::pos(g, x, y, z);
  float* out = F.out[AXIS];
  // clearance >= 2 at the cell above the face: both cells of the face are fluid
  if (SOLID && !(R.c >= 2.0f)) {
    int ox = AXIS == 0, oy = AXIS == 1, oz = AXIS == 2;
    if (solid_cell(g, solid, x, y, z) || solid_cell(g, solid, x - ox, y - oy, z - oz)) {
      stf(out, i, 0.0f);
      if (BOUNDS) {
        stf(F.lo[AXIS], i, __builtin_inff());
        stf(F.hi[AXIS], i, -__builtin_inff());
      }
      return;
    }
  }
  V3 vp = velocity_at_face<AXIS>(g, V, x, y, z, i);
  V3 mid = trace(p, -0.5f * dt, vp);
  if (SOLID) mid = clip_fast(g, solid, R, p, mid);
  V3 vm = velocity_at(g, V, mid);
  V3 dep = trace(p, -dt, vm);
  if (SOLID) dep = clip_fast(g, solid, R, p, dep);
  if (!BOUNDS) {
    stf(out, i, sample_face(F.in[AXIS], sx, sy, sz, FA::qx(g, dep), FA::qy(g, dep), FA::qz(g, dep)));
    return;
  }
  float lo, hi;
  stf(out, i, sample_face_mm(F.in[AXIS], sx, sy, sz, FA::qx(g, dep), FA::qy(g, dep), FA::qz(g, dep), lo, hi));
  const float lbx = FA::offx * g.hx, lby = FA::offy * g.hy, lbz = FA::offz * g.hz;
  const float hbx = (float(sx - 1) + FA::offx) * g.hx, hby = (float(sy - 1) + FA::offy) * g.hy,
              hbz = (float(sz - 1) + FA::offz) * g.hz;
  bool ok = !(dep.x < lbx || dep.y < lby || dep.z < lbz || dep.x > hbx || dep.y > hby || dep.z > hbz);
  if (SOLID && ok) ok = end_free(g, solid, clear, R, p, dep);
  stf(F.lo[AXIS], i, ok ? lo : __builtin_inff());
  stf(F.hi[AXIS], i, ok ? hi : -__builtin_inff());
}

// interior fast path of one face: departure point, then the stores
template <int AXIS>
__device__ __forceinline__ V3 face_mid_in(const GridDev& g, const Vel& V, float dt, int x, int y, int z, V3& p,
                                          uint32_t& i) {
  using FA = FaceAxis<AXIS>;
  i = FA::at(g, x, y, z);
  p = FA::pos(g, x, y, z);
  return trace(p, -0.5f * dt, velocity_at_face_in<AXIS>(g, V, x, y, z, i));
}

template <int AXIS, bool BOUNDS>
__device__ __forceinline__ void face_store_in(const GridDev& g, const FaceIO& F, uint32_t i, V3 dep) {
  using FA = FaceAxis<AXIS>;
  const float qx = FA::qx(g, dep), qy = FA::qy(g, dep), qz = FA::qz(g, dep);
  if (!BOUNDS) {
    stf(F.out[AXIS], i, sample_face_in(F.in[AXIS], FA::sx(g), FA::sy(g), qx, qy, qz));
    return;
  }
  float lo, hi;
  stf(F.out[AXIS], i, sample_face_mm_in(F.in[AXIS], FA::sx(g), FA::sy(g), qx, qy, qz, lo, hi));
  stf(F.lo[AXIS], i, lo);
  stf(F.hi[AXIS], i, hi);
}

template <bool SOLID, bool BOUNDS>
__global__ void k_face_predict(GridDev g, Vel V, float dt, FaceIO F, const uint8_t* __restrict__ solid,
                               const uint8_t* __restrict__ clear) {
  uint32_t ux, uy, uz, ue;
  if (!thread_site(g.nx + 1, g.ny + 1, ux, uy, uz, ue)) return;
  const int x = int(ux), y = int(uy), z = int(uz);
  const Reach R = SOLID ? reach_cell(g, clear, x, y, z) : Reach{-1.0f};
  if (__all(open_site<SOLID>(g, R, x, y, z))) {
    // an open site has all three faces
    V3 p0, p1, p2;
    uint32_t i0, i1, i2;
    const V3 m0 = face_mid_in<0>(g, V, dt, x, y, z, p0, i0), m1 = face_mid_in<1>(g, V, dt, x, y, z, p1, i1),
             m2 = face_mid_in<2>(g, V, dt, x, y, z, p2, i2);
    if (__all(near_site(g, p0, m0) && near_site(g, p1, m1) && near_site(g, p2, m2))) {
      const V3 d0 = trace(p0, -dt, velocity_at_in(g, V, m0)), d1 = trace(p1, -dt, velocity_at_in(g, V, m1)),
               d2 = trace(p2, -dt, velocity_at_in(g, V, m2));
      if (__all(near_site(g, p0, d0) && near_site(g, p1, d1) && near_site(g, p2, d2))) {
        face_store_in<0, BOUNDS>(g, F, i0, d0);
        face_store_in<1, BOUNDS>(g, F, i1, d1);
        face_store_in<2, BOUNDS>(g, F, i2, d2);
        return;
      }
    }
  }
  if (FaceAxis<0>::exists(g, x, y, z)) face_predict_general<0, SOLID, BOUNDS>(g, V, dt, F, solid, clear, R, x, y, z);
  if (FaceAxis<1>::exists(g, x, y, z)) face_predict_general<1, SOLID, BOUNDS>(g, V, dt, F, solid, clear, R, x, y, z);
  if (FaceAxis<2>::exists(g, x, y, z)) face_predict_general<2, SOLID, BOUNDS>(g, V, dt, F, solid, clear, R, x, y, z);
}

// ---------------------------------------------------------------------------
// fused velocity transport + jet relaxation + buoyancy (W)
// ---------------------------------------------------------------------------
struct VelArgs {
  GridDev g;
  Vel V;
  float dt;
  const float* pred[3];
  float* out[3];
  const float* density;
  const float* temperature;
  float lift, weight;
  const uint8_t* solid;
  const uint8_t* clear;
  const float* lo[3];  // donor bounds from the predictor pass (BOUNDS variant)
  const float* hi[3];
};

// Transported value of one face before the jet relaxation and buoyancy
// (general path).  BOUNDS (MacCormack only): the predictor pass already
// sampled the old field at the departure point and stored the donor bounds,
// so only the forward trace and the reverse sample remain here.  blocked: a
// face with a solid cell on either side (value 0, no relaxation).
template <int AXIS, bool SOLID, bool MC, bool BOUNDS>
__device__ float velocity_general(const VelArgs& A, const Reach& R, int x, int y, int z, bool& blocked) {
  using FA = FaceAxis<AXIS>;
  const GridDev& g = A.g;
  const int sx = FA::sx(g), sy = FA::sy(g), sz = FA::sz(g);
  const uint32_t i = FA::at(g, x, y, z);
  const float* orig = AXIS == 0 ? A.V.u : (AXIS == 1 ? A.V.v : A.V.w);
  const float* pred = A.pred[AXIS];
  const float offx = FA::offx, offy = FA::offy, offz = FA::offz;
  const float hx = g.hx, hy = g.hy, hz = g.hz;
  const float dt = A.dt;
  const V3 p = FA::pos(g, x, y, z);
  float value = 0.0f;
  blocked = false;
  // clearance >= 2 at the cell above the face: both cells of the face are fluid
  if (SOLID && !(R.c >= 2.0f)) {
    int ox = AXIS == 0, oy = AXIS == 1, oz = AXIS == 2;
    blocked = solid_cell(g, A.solid, x, y, z) || solid_cell(g, A.solid, x - ox, y - oy, z - oz);
  }
  if (blocked) return 0.0f;
  if (BOUNDS) {
    V3 vp = velocity_at_face<AXIS>(g, A.V, x, y, z, i);
    const float low = ldf(pred, i), lo = ldf(A.lo[AXIS], i), hi = ldf(A.hi[AXIS], i);
    value = low;
    if (lo <= hi) {
      const float hbx = (float(sx - 1) + offx) * hx, hby = (float(sy - 1) + offy) * hy,
                  hbz = (float(sz - 1) + offz) * hz;
      V3 rmid = trace(p, 0.5f * dt, vp);
      V3 va = velocity_at(g, A.V, rmid);
      V3 arr = trace(p, dt, va);
      bool ok = !(arr.x < offx * hx || arr.y < offy * hy || arr.z < offz * hz || arr.x > hbx || arr.y > hby ||
                  arr.z > hbz);
      if (SOLID && ok && !(R.free(g, p, rmid) && R.free(g, p, arr)))
        ok = !halo_cell(g, A.solid, A.clear, x, y, z) && !halo_fast(g, A.solid, A.clear, arr) &&
             dist3(clip_fast(g, A.solid, R, p, rmid), rmid) < 1e-7f &&
             dist3(clip_fast(g, A.solid, R, p, arr), arr) < 1e-7f;
      if (ok) {
        float rev = sample_face(pred, sx, sy, sz, FA::qx(g, arr), FA::qy(g, arr), FA::qz(g, arr));
        float corr = __builtin_fmaf(0.5f, ldf(orig, i) - rev, ldf(pred, i));
        value = __builtin_fmaf(0.5f, low, 0.5f * fminf(fmaxf(corr, lo), hi));
      }
    }
    return value;
  }
  V3 vp = velocity_at_face<AXIS>(g, A.V, x, y, z, i);
  V3 mid = trace(p, -0.5f * dt, vp);
  if (SOLID) mid = clip_fast(g, A.solid, R, p, mid);
  V3 vm = velocity_at(g, A.V, mid);
  V3 dep = trace(p, -dt, vm);
  if (SOLID) dep = clip_fast(g, A.solid, R, p, dep);
  float qx = FA::qx(g, dep), qy = FA::qy(g, dep), qz = FA::qz(g, dep);
  float lo = 0.0f, hi = 0.0f;
  float low = MC ? sample_face_mm(orig, sx, sy, sz, qx, qy, qz, lo, hi) : sample_face(orig, sx, sy, sz, qx, qy, qz);
  value = low;
  if (MC) {
    float lbx = offx * hx, lby = offy * hy, lbz = offz * hz;
    float hbx = (float(sx - 1) + offx) * hx, hby = (float(sy - 1) + offy) * hy, hbz = (float(sz - 1) + offz) * hz;
    V3 rmid = trace(p, 0.5f * dt, vp);
    V3 va = velocity_at(g, A.V, rmid);
    V3 arr = trace(p, dt, va);
    bool ok = true;
    if (SOLID) ok = safe_fast(g, A.solid, A.clear, R, p, x, y, z, dep, rmid, arr);
    ok = ok && !(dep.x < lbx || dep.y < lby || dep.z < lbz || dep.x > hbx || dep.y > hby || dep.z > hbz);
    ok = ok && !(arr.x < lbx || arr.y < lby || arr.z < lbz || arr.x > hbx || arr.y > hby || arr.z > hbz);
    if (ok) {
      // dep lies inside the face lattice here, so the sampled donors are the clamp set
      float rev = sample_face(pred, sx, sy, sz, FA::qx(g, arr), FA::qy(g, arr), FA::qz(g, arr));
      float corr = __builtin_fmaf(0.5f, ldf(orig, i) - rev, ldf(pred, i));
      value = __builtin_fmaf(0.5f, low, 0.5f * fminf(fmaxf(corr, lo), hi));
    }
  }
  return value;
}

// interior fast path of one face: the forward trace's midpoint (BOUNDS) or
// the back-trace's (semi-Lagrangian)
template <int AXIS, bool BOUNDS>
__device__ __forceinline__ V3 velocity_mid_in(const VelArgs& A, int x, int y, int z, V3& p, uint32_t& i) {
  using FA = FaceAxis<AXIS>;
  i = FA::at(A.g, x, y, z);
  p = FA::pos(A.g, x, y, z);
  return trace(p, (BOUNDS ? 0.5f : -0.5f) * A.dt, velocity_at_face_in<AXIS>(A.g, A.V, x, y, z, i));
}

// value of one face from its interior trace end t (arrival for BOUNDS,
// departure for semi-Lagrangian)
template <int AXIS, bool BOUNDS>
__device__ __forceinline__ float velocity_value_in(const VelArgs& A, uint32_t i, V3 t) {
  using FA = FaceAxis<AXIS>;
  const GridDev& g = A.g;
  const float* orig = AXIS == 0 ? A.V.u : (AXIS == 1 ? A.V.v : A.V.w);
  if (!BOUNDS) return sample_face_in(orig, FA::sx(g), FA::sy(g), FA::qx(g, t), FA::qy(g, t), FA::qz(g, t));
  const float* pred = A.pred[AXIS];
  const float low = ldf(pred, i), lo = ldf(A.lo[AXIS], i), hi = ldf(A.hi[AXIS], i);
  float value = low;
  if (lo <= hi) {
    const float rev = sample_face_in(pred, FA::sx(g), FA::sy(g), FA::qx(g, t), FA::qy(g, t), FA::qz(g, t));
    const float corr = __builtin_fmaf(0.5f, ldf(orig, i) - rev, ldf(pred, i));
    value = __builtin_fmaf(0.5f, low, 0.5f * fminf(fmaxf(corr, lo), hi));
  }
  return value;
}

// jet relaxation (unblocked faces) and buoyancy (W), then the store
template <int AXIS>
__device__ __forceinline__ void velocity_finish(const VelArgs& A, const fx_sources& src, int x, int y, int z,
                                                float value, bool blocked) {
  using FA = FaceAxis<AXIS>;
  const GridDev& g = A.g;
  const float dt = A.dt;
  if (!blocked) {
    const V3 p = FA::pos(g, x, y, z);
    // jet relaxation: coupling rates summed first (order independent)
    float total = 0.0f, weighted = 0.0f;
    for (int s = 0; s < src.count; ++s) {
      const float* row = src.table + 16 * s;
      float rate = row[7] * emission_weight(p, row);
      total += rate;
      weighted += rate * row[8 + AXIS];
    }
    if (total > 0.0f) {
      float blend = -expm1f(-dt * total);
      value = __builtin_fmaf(blend, weighted / total - value, value);
    }
  }
  if (AXIS == 2 && A.density) {
    int zb = min(max(z - 1, 0), g.nz - 1), za = min(max(z, 0), g.nz - 1);
    const uint32_t ib = lin(x, y, zb, g.nx, g.ny) << 2, ia = lin(x, y, za, g.nx, g.ny) << 2;
    float rho = 0.5f * (ldf(A.density, ib) + ldf(A.density, ia));
    float theta = 0.5f * (ldf(A.temperature, ib) + ldf(A.temperature, ia));
    value = __builtin_fmaf(dt, A.lift * theta - A.weight * rho, value);
  }
  stf(A.out[AXIS], FA::at(g, x, y, z), value);
}

template <bool SOLID, bool MC, bool BOUNDS>
__global__ void k_velocity(VelArgs A, fx_sources src) {
  const GridDev& g = A.g;
  uint32_t ux, uy, uz, ue;
  if (!thread_site(g.nx + 1, g.ny + 1, ux, uy, uz, ue)) return;
  const int x = int(ux), y = int(uy), z = int(uz);
  const Reach R = SOLID ? reach_cell(g, A.clear, x, y, z) : Reach{-1.0f};
  if ((BOUNDS || !MC) && __all(open_site<SOLID>(g, R, x, y, z))) {
    V3 p0, p1, p2;
    uint32_t i0, i1, i2;
    const V3 m0 = velocity_mid_in<0, BOUNDS>(A, x, y, z, p0, i0), m1 = velocity_mid_in<1, BOUNDS>(A, x, y, z, p1, i1),
             m2 = velocity_mid_in<2, BOUNDS>(A, x, y, z, p2, i2);
    if (__all(near_site(g, p0, m0) && near_site(g, p1, m1) && near_site(g, p2, m2))) {
      const float s = BOUNDS ? A.dt : -A.dt;
      const V3 t0 = trace(p0, s, velocity_at_in(g, A.V, m0)), t1 = trace(p1, s, velocity_at_in(g, A.V, m1)),
               t2 = trace(p2, s, velocity_at_in(g, A.V, m2));
      if (__all(near_site(g, p0, t0) && near_site(g, p1, t1) && near_site(g, p2, t2))) {
        const float v0 = velocity_value_in<0, BOUNDS>(A, i0, t0), v1 = velocity_value_in<1, BOUNDS>(A, i1, t1),
                    v2 = velocity_value_in<2, BOUNDS>(A, i2, t2);
        velocity_finish<0>(A, src, x, y, z, v0, false);
        velocity_finish<1>(A, src, x, y, z, v1, false);
        velocity_finish<2>(A, src, x, y, z, v2, false);
        return;
      }
    }
  }
  bool blocked;
  if (FaceAxis<0>::exists(g, x, y, z)) {
    const float v = velocity_general<0, SOLID, MC, BOUNDS>(A, R, x, y, z, blocked);
    velocity_finish<0>(A, src, x, y, z, v, blocked);
  }
  if (FaceAxis<1>::exists(g, x, y, z)) {
    const float v = velocity_general<1, SOLID, MC, BOUNDS>(A, R, x, y, z, blocked);
    velocity_finish<1>(A, src, x, y, z, v, blocked);
  }
  if (FaceAxis<2>::exists(g, x, y, z)) {
    const float v = velocity_general<2, SOLID, MC, BOUNDS>(A, R, x, y, z, blocked);
    velocity_finish<2>(A, src, x, y, z, v, blocked);
  }
}

// ---------------------------------------------------------------------------
// scalar sampling (closed box: integer indices clamped)
// ---------------------------------------------------------------------------
// Corner k = (k & 1, k >> 1 & 1, k >> 2) of a cell-array sample: element
// offset of the low corner plus element steps to the high corners along
// x/y/z (0 where the high index clamps onto the low one).
struct Corners {
  uint32_t o, dx, dy, dz;
  float w[8];
  float norm;  // 1 / sum of fluid weights (solid sampling)
  __device__ __forceinline__ uint32_t at(int k) const {
    return o + ((k & 1) ? dx : 0u) + ((k & 2) ? dy : 0u) + ((k & 4) ? dz : 0u);
  }
};

// zero_ext: corners outside the grid contribute 0 (shaders/mac_advect_density.glsl
// fetchZero); otherwise integer indices are clamped (shaders/closed_sample.glsl).
__device__ __forceinline__ void closed_corners(const GridDev& g, float qx, float qy, float qz, Corners& c,
                                               bool zero_ext = false) {
  const float lx = floorf(qx), ly = floorf(qy), lz = floorf(qz);
  const int bx = int(lx), by = int(ly), bz = int(lz);
  const float fx = qx - lx, fy = qy - ly, fz = qz - lz;
  const int x0 = min(max(bx, 0), g.nx - 1), x1 = min(max(bx + 1, 0), g.nx - 1);
  const int y0 = min(max(by, 0), g.ny - 1), y1 = min(max(by + 1, 0), g.ny - 1);
  const int z0 = min(max(bz, 0), g.nz - 1), z1 = min(max(bz + 1, 0), g.nz - 1);
  c.o = lin(x0, y0, z0, g.nx, g.ny);
  c.dx = x1 != x0 ? 1u : 0u;
  c.dy = y1 != y0 ? uint32_t(g.nx) : 0u;
  c.dz = z1 != z0 ? mul24(uint32_t(g.nx), uint32_t(g.ny)) : 0u;
  // per-axis weights of the low/high corner; an out-of-range corner weighs 0
  float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
  if (zero_ext) {
    if (x0 != bx) wx[0] = 0.0f;
    if (x1 != bx + 1) wx[1] = 0.0f;
    if (y0 != by) wy[0] = 0.0f;
    if (y1 != by + 1) wy[1] = 0.0f;
    if (z0 != bz) wz[0] = 0.0f;
    if (z1 != bz + 1) wz[1] = 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) c.w[k] = wx[k & 1] * wy[(k >> 1) & 1] * wz[k >> 2];
}

// Solid corners drop out and the rest renormalise (shaders/closed_sample.glsl:9-19):
// done once per corner set, then every channel is a plain weighted sum.
template <bool SOLID>
__device__ __forceinline__ void fluid_weights(Corners& c, const uint8_t* __restrict__ solid) {
  if (!SOLID) return;
  float t = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (solid[c.at(k)]) c.w[k] = 0.0f;
    t += c.w[k];
  }
  c.norm = t > 1e-8f ? 1.0f / t : 0.0f;
}

template <bool SOLID>
__device__ __forceinline__ float closed_sample(const float* __restrict__ f, const Corners& c) {
  float v = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) v = __builtin_fmaf(c.w[k], ldf(f, c.at(k) << 2), v);
  if (!SOLID) return v;
  return c.norm > 0.0f ? v * c.norm : 0.0f;
}

// closed_corners for q at least 1 from both ends of every axis (no clamps;
// no corner out of range, so zero extension changes nothing)
__device__ __forceinline__ void closed_corners_in(const GridDev& g, float qx, float qy, float qz, Corners& c) {
  const float lx = floorf(qx), ly = floorf(qy), lz = floorf(qz);
  const float fx = qx - lx, fy = qy - ly, fz = qz - lz;
  c.o = lin(uint32_t(lx), uint32_t(ly), uint32_t(lz), g.nx, g.ny);
  c.dx = 1u;
  c.dy = uint32_t(g.nx);
  c.dz = mul24(uint32_t(g.nx), uint32_t(g.ny));
  float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
#pragma unroll
  for (int k = 0; k < 8; ++k) c.w[k] = wx[k & 1] * wy[(k >> 1) & 1] * wz[k >> 2];
}

// the eight corner values of a corner set, in corner order
__device__ __forceinline__ void corner_values(const float* __restrict__ f, const Corners& c, float (&d)[8]) {
#pragma unroll
  for (int k = 0; k < 8; ++k) d[k] = ldf(f, c.at(k) << 2);
}

// closed_sample on loaded corner values
template <bool SOLID>
__device__ __forceinline__ float weighted_sample(const Corners& c, const float (&d)[8]) {
  float v = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) v = __builtin_fmaf(c.w[k], d[k], v);
  if (!SOLID) return v;
  return c.norm > 0.0f ? v * c.norm : 0.0f;
}

// fluid_weights when all eight corners are fluid
template <bool SOLID>
__device__ __forceinline__ void fluid_norm(Corners& c) {
  if (!SOLID) return;
  float t = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) t += c.w[k];
  c.norm = t > 1e-8f ? 1.0f / t : 0.0f;
}

struct ScalarSet {
  const float* f[3];
  float* o[3];
  float* lo[3];  // donor bounds (BOUNDS variant)
  float* hi[3];
  BrickMask bricks;  // computed bricks (the predictor has no zero-store state)
};

// BOUNDS: also store, per channel, the min/max of the eight donors of the
// departure sample when the departure point may be MacCormack-corrected, else
// lo = +inf, hi = -inf
template <bool SOLID, bool BOUNDS = false>
__global__ void k_scalar_predict(GridDev g, Vel V, float dt, int channels, ScalarSet S_,
                                 const uint8_t* __restrict__ solid, const uint8_t* __restrict__ clear) {
  uint32_t ux, uy, uz, ue;
  if (!thread_site(g.nx, g.ny, ux, uy, uz, ue)) return;
  const int x = int(ux), y = int(uy), z = int(uz);
  if (S_.bricks.state(x, y, z) != kBrickCompute) return;
  const uint32_t i = ue << 2;  // byte offset of this cell
  const float hx = g.hx, hy = g.hy, hz = g.hz;
  V3 p = v3((x + 0.5f) * hx, (y + 0.5f) * hy, (z + 0.5f) * hz);
  const Reach R = SOLID ? reach_cell(g, clear, x, y, z) : Reach{-1.0f};
  if (__all(open_site<SOLID>(g, R, x, y, z))) {
    const V3 vp = velocity_at_cell(g, V, x, y, z);
    const V3 mid = trace(p, -0.5f * dt, vp);
    if (__all(near_site(g, p, mid))) {
      const V3 vm = velocity_at_in(g, V, mid);
      const V3 dep = trace(p, -dt, vm);
      if (__all(near_site(g, p, dep))) {
        Corners c;
        closed_corners_in(g, lat(dep.x, g.ix, 0.5f), lat(dep.y, g.iy, 0.5f), lat(dep.z, g.iz, 0.5f), c);
        fluid_norm<SOLID>(c);
        for (int k = 0; k < channels; ++k) {
          float d[8];
          corner_values(S_.f[k], c, d);
          stf(S_.o[k], i, weighted_sample<SOLID>(c, d));
          if (BOUNDS) {
            float lo = __builtin_inff(), hi = -__builtin_inff();
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              lo = fminf(lo, d[q]);
              hi = fmaxf(hi, d[q]);
            }
            stf(S_.lo[k], i, lo);
            stf(S_.hi[k], i, hi);
          }
        }
        return;
      }
    }
  }
  // clearance >= 1: the cell is fluid
  if (SOLID && !(R.c >= 1.0f) && solid[ue]) {
    for (int k = 0; k < channels; ++k) {
      stf(S_.o[k], i, 0.0f);
      if (BOUNDS) {
        stf(S_.lo[k], i, __builtin_inff());
        stf(S_.hi[k], i, -__builtin_inff());
      }
    }
    return;
  }
  V3 vp = velocity_at_cell(g, V, x, y, z);
  V3 mid = trace(p, -0.5f * dt, vp);
  if (SOLID) mid = clip_fast(g, solid, R, p, mid);
  V3 vm = velocity_at(g, V, mid);
  V3 dep = trace(p, -dt, vm);
  if (SOLID) dep = clip_fast(g, solid, R, p, dep);
  Corners c;
  closed_corners(g, lat(dep.x, g.ix, 0.5f), lat(dep.y, g.iy, 0.5f), lat(dep.z, g.iz, 0.5f), c);
  fluid_weights<SOLID>(c, solid);  // the donors (c.at) stay all eight corners
  bool ok = true;
  if (BOUNDS) {
    const float lx = 0.5f * hx, ly = 0.5f * hy, lz = 0.5f * hz;
    const float ux = g.ex - 0.5f * hx, uy = g.ey - 0.5f * hy, uz = g.ez - 0.5f * hz;
    ok = !(dep.x < lx || dep.y < ly || dep.z < lz || dep.x > ux || dep.y > uy || dep.z > uz);
    if (SOLID && ok) ok = end_free(g, solid, clear, R, p, dep);
  }
  for (int k = 0; k < channels; ++k) {
    const float* f = S_.f[k];
    stf(S_.o[k], i, closed_sample<SOLID>(f, c));
    if (BOUNDS) {
      float lo = __builtin_inff(), hi = -__builtin_inff();
      if (ok) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const float d = ldf(f, c.at(q) << 2);
          lo = fminf(lo, d);
          hi = fmaxf(hi, d);
        }
      }
      stf(S_.lo[k], i, lo);
      stf(S_.hi[k], i, hi);
    }
  }
}

struct ScalarArgs {
  GridDev g;
  Vel V;
  float dt;
  int channels;
  const float* f[3];
  const float* pred[3];
  float* out[3];
  float keep[3];  // exp(-decay*dt)
  int chan[3];
  int zero_ext;
  const uint8_t* solid;
  const uint8_t* clear;
  const float* lo[3];  // donor bounds from the predictor pass (BOUNDS variant)
  const float* hi[3];
  BrickMask bricks;    // plan: skip, compute or store 0 per brick
  uint8_t* live[3];    // nullable: byte of the site's brick set to 1 when the channel's output is not 0
};

// the channel's output is nonzero (or NaN): mark its brick live
__device__ __forceinline__ void mark_live(uint8_t* live, const BrickMask& m, int x, int y, int z, float v) {
  if (live && !(v == 0.0f)) live[m.at(x, y, z)] = 1;
}

// BOUNDS (MacCormack only): the predictor pass already holds the
// semi-Lagrangian value and the donor bounds of every channel, so only the
// forward trace and the reverse samples remain here.
template <bool SOLID, bool MC, bool BOUNDS = false>
__global__ void k_scalar_advect(ScalarArgs A, fx_sources src) {
  const GridDev& g = A.g;
  uint32_t ux, uy, uz, ue;
  if (!thread_site(g.nx, g.ny, ux, uy, uz, ue)) return;
  const int x = int(ux), y = int(uy), z = int(uz);
  const uint8_t plan = A.bricks.state(x, y, z);
  if (plan == kBrickSkip) return;
  const uint32_t i = ue << 2;  // byte offset of this cell
  if (plan == kBrickZero) {
    for (int k = 0; k < A.channels; ++k) stf(A.out[k], i, 0.0f);
    return;
  }
  const float dt = A.dt;
  const float hx = g.hx, hy = g.hy, hz = g.hz;
  V3 p = v3((x + 0.5f) * hx, (y + 0.5f) * hy, (z + 0.5f) * hz);
  const Reach R = SOLID ? reach_cell(g, A.clear, x, y, z) : Reach{-1.0f};
  // clearance >= 1: the cell is fluid
  if (SOLID && !(R.c >= 1.0f) && A.solid[ue]) {
    for (int k = 0; k < A.channels; ++k) stf(A.out[k], i, 0.0f);
    return;
  }
  V3 vp = velocity_at_cell(g, A.V, x, y, z);
  float wsrc[8];
  for (int s = 0; s < src.count; ++s) wsrc[s] = emission_weight(p, src.table + 16 * s);
  if (BOUNDS) {
    // forward trace and its fluid/clip conditions (the departure ones are in lo <= hi)
    V3 rmid = trace(p, 0.5f * dt, vp);
    // interior fast path (open site, trace within 2 cells): same values
    bool fast = __all(open_site<SOLID>(g, R, x, y, z)) && __all(near_site(g, p, rmid));
    V3 va = fast ? velocity_at_in(g, A.V, rmid) : velocity_at(g, A.V, rmid);
    V3 arr = trace(p, dt, va);
    fast = fast && __all(near_site(g, p, arr));
    bool ok_arr = true;
    Corners ca;
    if (fast) {
      closed_corners_in(g, lat(arr.x, g.ix, 0.5f), lat(arr.y, g.iy, 0.5f), lat(arr.z, g.iz, 0.5f), ca);
      fluid_norm<SOLID>(ca);
    } else {
      const float ux = g.ex - 0.5f * hx, uy = g.ey - 0.5f * hy, uz = g.ez - 0.5f * hz;
      ok_arr = !(arr.x < 0.5f * hx || arr.y < 0.5f * hy || arr.z < 0.5f * hz || arr.x > ux || arr.y > uy ||
                 arr.z > uz);
      if (SOLID && ok_arr && !(R.free(g, p, rmid) && R.free(g, p, arr)))
        ok_arr = !halo_cell(g, A.solid, A.clear, x, y, z) && !halo_fast(g, A.solid, A.clear, arr) &&
                 dist3(clip_fast(g, A.solid, R, p, rmid), rmid) < 1e-7f &&
                 dist3(clip_fast(g, A.solid, R, p, arr), arr) < 1e-7f;
      closed_corners(g, lat(arr.x, g.ix, 0.5f), lat(arr.y, g.iy, 0.5f), lat(arr.z, g.iz, 0.5f), ca);
      fluid_weights<SOLID>(ca, A.solid);
    }
    for (int k = 0; k < A.channels; ++k) {
      const float low = ldf(A.pred[k], i), lo = ldf(A.lo[k], i), hi = ldf(A.hi[k], i);
      float value = low;
      if (ok_arr && lo <= hi) {
        float rev;
        if (fast) {
          float d[8];
          corner_values(A.pred[k], ca, d);
          rev = weighted_sample<SOLID>(ca, d);
        } else {
          rev = closed_sample<SOLID>(A.pred[k], ca);
        }
        float corr = __builtin_fmaf(0.5f, ldf(A.f[k], i) - rev, low);
        value = fminf(fmaxf(corr, lo), hi);
      }
      value *= A.keep[k];
      int ch = A.chan[k];
      float target = 0.0f;
      bool any_target = false;
      for (int s = 0; s < src.count; ++s) {
        const float* row = src.table + 16 * s;
        float rate = ch == 0 ? row[11] : (ch == 1 ? row[12] : row[15]);
        if (ch == 0 && row[13] > 0.5f) {
          target = fmaxf(target, rate * wsrc[s]);
          any_target = true;
        } else {
          value = __builtin_fmaf(dt * rate, wsrc[s], value);
        }
      }
      if (ch == 0 && (src.multi || any_target)) value = fmaxf(value, target);
      stf(A.out[k], i, value);
      mark_live(A.live[k], A.bricks, x, y, z, value);
    }
    return;
  }
  V3 mid = trace(p, -0.5f * dt, vp);
  if (SOLID) mid = clip_fast(g, A.solid, R, p, mid);
  V3 vm = velocity_at(g, A.V, mid);
  V3 dep = trace(p, -dt, vm);
  if (SOLID) dep = clip_fast(g, A.solid, R, p, dep);
  Corners c;
  closed_corners(g, lat(dep.x, g.ix, 0.5f), lat(dep.y, g.iy, 0.5f), lat(dep.z, g.iz, 0.5f), c, A.zero_ext != 0);
  // min/max donors: all eight corners of the departure cell (c.at), solid or not
  fluid_weights<SOLID>(c, A.solid);
  bool ok = false;
  Corners ca;
  if (MC) {
    V3 rmid = trace(p, 0.5f * dt, vp);
    V3 va = velocity_at(g, A.V, rmid);
    V3 arr = trace(p, dt, va);
    ok = true;
    if (SOLID) ok = safe_fast(g, A.solid, A.clear, R, p, x, y, z, dep, rmid, arr);
    float lx = 0.5f * hx, ly = 0.5f * hy, lz = 0.5f * hz;
    float ux = g.ex - 0.5f * hx, uy = g.ey - 0.5f * hy, uz = g.ez - 0.5f * hz;
    ok = ok && !(dep.x < lx || dep.y < ly || dep.z < lz || dep.x > ux || dep.y > uy || dep.z > uz);
    ok = ok && !(arr.x < lx || arr.y < ly || arr.z < lz || arr.x > ux || arr.y > uy || arr.z > uz);
    closed_corners(g, lat(arr.x, g.ix, 0.5f), lat(arr.y, g.iy, 0.5f), lat(arr.z, g.iz, 0.5f), ca);
    fluid_weights<SOLID>(ca, A.solid);
  }
  for (int k = 0; k < A.channels; ++k) {
    const float* f = A.f[k];
    float value = closed_sample<SOLID>(f, c);
    if (MC && ok) {
      float rev = closed_sample<SOLID>(A.pred[k], ca);
      float corr = __builtin_fmaf(0.5f, ldf(f, i) - rev, ldf(A.pred[k], i));
      float lo = 3.4e38f, hi = -3.4e38f;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float d = ldf(f, c.at(q) << 2);
        lo = fminf(lo, d);
        hi = fmaxf(hi, d);
      }
      value = fminf(fmaxf(corr, lo), hi);
    }
    value *= A.keep[k];
    int ch = A.chan[k];
    float target = 0.0f;
    bool any_target = false;
    for (int s = 0; s < src.count; ++s) {
      const float* row = src.table + 16 * s;
      float rate = ch == 0 ? row[11] : (ch == 1 ? row[12] : row[15]);
      if (ch == 0 && row[13] > 0.5f) {
        target = fmaxf(target, rate * wsrc[s]);
        any_target = true;
      } else {
        value = __builtin_fmaf(dt * rate, wsrc[s], value);
      }
    }
    if (ch == 0 && (src.multi || any_target)) value = fmaxf(value, target);
    stf(A.out[k], i, value);
    mark_live(A.live[k], A.bricks, x, y, z, value);
  }
}

// ---------------------------------------------------------------------------
__global__ void k_combustion(int64_t n, float dt, float ign, float k, float heat, float smoke,
                             const float* __restrict__ fuel, const float* __restrict__ T,
                             const float* __restrict__ rho, float* __restrict__ fo, float* __restrict__ To,
                             float* __restrict__ ro, float* __restrict__ flame) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f = fuel[i], t = T[i];
  float burned = t >= ign ? fmaxf(f, 0.0f) * -expm1f(-k * dt) : 0.0f;
  float fl = burned / dt;
  fo[i] = fmaxf(f - dt * fl, 0.0f);
  To[i] = t + dt * heat * fl;
  ro[i] = rho[i] + dt * smoke * fl;
  flame[i] = fl;
}

// k_combustion over the bricks of a plan (skipped bricks untouched, zero
// bricks stored 0); live[0..3]: fuel, temperature, density and flame outputs
__global__ void k_combustion_active(int nx, int ny, int n, float dt, float ign, float k, float heat, float smoke,
                                    const float* __restrict__ fuel, const float* __restrict__ T,
                                    const float* __restrict__ rho, float* __restrict__ fo, float* __restrict__ To,
                                    float* __restrict__ ro, float* __restrict__ flame, BrickMask m,
                                    uint8_t* lf, uint8_t* lt, uint8_t* lr, uint8_t* ll) {
  const int i = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
  if (i >= n) return;
  const int row = i / nx, x = i - row * nx, y = row % ny, z = row / ny;
  const uint8_t plan = m.state(x, y, z);
  if (plan == kBrickSkip) return;
  if (plan == kBrickZero) {
    fo[i] = To[i] = ro[i] = flame[i] = 0.0f;
    return;
  }
  float f = fuel[i], t = T[i];
  float burned = t >= ign ? fmaxf(f, 0.0f) * -expm1f(-k * dt) : 0.0f;
  float fl = burned / dt;
  const float a = fmaxf(f - dt * fl, 0.0f), b = t + dt * heat * fl, c = rho[i] + dt * smoke * fl;
  fo[i] = a;
  To[i] = b;
  ro[i] = c;
  flame[i] = fl;
  mark_live(lf, m, x, y, z, a);
  mark_live(lt, m, x, y, z, b);
  mark_live(lr, m, x, y, z, c);
  mark_live(ll, m, x, y, z, fl);
}

struct MaxArgs {
  const float* a[8];
  int64_t n[8];
};

__global__ void k_max_init(float* out, int count) {
  if (int(threadIdx.x) < count) out[threadIdx.x] = 0.0f;
}

// 16-byte loads, four in flight per thread, on a 16-byte aligned array (the
// tail and unaligned arrays element by element); the maximum does not depend
// on the order
__global__ void k_max_abs(MaxArgs A, float* out) {
  const int k = blockIdx.y;
  const float* a = A.a[k];
  const int64_t n = A.n[k];
  float m = 0.0f;
  auto acc = [&](float v) { m = fmaxf(m, isfinite(v) ? fabsf(v) : 3.4e38f); };
  auto acc4 = [&](const float4& v) {
    acc(v.x);
    acc(v.y);
    acc(v.z);
    acc(v.w);
  };
  const int64_t nq = (reinterpret_cast<uintptr_t>(a) & 15u) ? 0 : n / 4;
  const float4* a4 = reinterpret_cast<const float4*>(a);
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; q + 3 * stride < nq; q += 4 * stride) {
    const float4 v0 = a4[q], v1 = a4[q + stride], v2 = a4[q + 2 * stride], v3 = a4[q + 3 * stride];
    acc4(v0);
    acc4(v1);
    acc4(v2);
    acc4(v3);
  }
  for (; q < nq; q += stride) acc4(a4[q]);
  for (int64_t i = 4 * nq + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) acc(a[i]);
  m = wave_max(m);
  __shared__ float part[4];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float r = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
    atomicMax(reinterpret_cast<unsigned int*>(out + k), __float_as_uint(r));
  }
}

static const dim3 kBlock(64, 4, 1);

// the clearance of the cells in box b = [x0, x1) x [y0, y1) x [z0, z1) from
// the solid mask, the other cells' kept: a cell's clearance depends on the
// mask within kClearanceRadius cells, so pass 0 runs over the box widened by
// the radius along y and z, pass 1 along z (both into scratch arrays, as the
// widened cells' own clearance must not change), pass 2 over the box
int launch_clearance_box(const fx_grid& g, const int* b, const uint8_t* solid, uint8_t* clearance, uint8_t* scratch,
                         uint8_t* scratch2, hipStream_t s) {
  const int r = kClearanceRadius;
  const int y0 = max(b[1] - r, 0), y1 = min(b[4] + r, g.ny), z0 = max(b[2] - r, 0), z1 = min(b[5] + r, g.nz);
  const int bx = b[3] - b[0], by = b[4] - b[1], bz = b[5] - b[2];
  if (bx <= 0 || by <= 0 || bz <= 0) return FX_OK;
  auto grid = [](int ex, int ey, int ez) { return dim3((ex + 63) / 64, (ey + 3) / 4, ez); };
  hipLaunchKernelGGL(k_clearance, grid(bx, y1 - y0, z1 - z0), kBlock, 0, s, g.nx, g.ny, g.nz, 0, solid, scratch, r,
                     b[0], y0, z0, b[3], y1);
  hipLaunchKernelGGL(k_clearance, grid(bx, by, z1 - z0), kBlock, 0, s, g.nx, g.ny, g.nz, 1, (const uint8_t*)scratch,
                     scratch2, r, b[0], b[1], z0, b[3], b[4]);
  hipLaunchKernelGGL(k_clearance, grid(bx, by, bz), kBlock, 0, s, g.nx, g.ny, g.nz, 2, (const uint8_t*)scratch2,
                     clearance, r, b[0], b[1], b[2], b[3], b[4]);
  FX_LAUNCHED("clearance box");
  return FX_OK;
}

static std::atomic<int> g_fast_paths{1};

GridDev dev_grid(const fx_grid& g) {
  return GridDev{g.nx,        g.ny,        g.nz,        g.ex,        g.ey,        g.ez,
                 g.ex / g.nx, g.ey / g.ny, g.ez / g.nz, g.nx / g.ex, g.ny / g.ey, g.nz / g.ez,
                 g_fast_paths.load()};
}

// transport kernels: 256 threads per block, 16 x 16 sites of one z-plane of an
// sx x sy x sz array per block (thread_site)
static const dim3 kSiteBlock(256);
static dim3 site_grid(int sx, int sy, int sz) {
  return dim3(((unsigned(sx) + kSiteTile - 1) / kSiteTile) * ((unsigned(sy) + kSiteTile - 1) / kSiteTile), sz);
}

// 32-bit byte offsets and 24-bit products of the transport kernels: every
// face array below 2^30 elements, every face plane below 2^24
int check_transport_grid(const fx_grid& g) {
  if (int e = check_grid(g)) return e;
  const int64_t nx = g.nx, ny = g.ny, nz = g.nz;
  FX_REQUIRE((nx + 1) * ny * nz < (int64_t(1) << 30) && nx * (ny + 1) * nz < (int64_t(1) << 30) &&
                 nx * ny * (nz + 1) < (int64_t(1) << 30),
             FX_EINVAL, "transport arrays are limited to 2^30 elements (grid %dx%dx%d)", g.nx, g.ny, g.nz);
  FX_REQUIRE((nx + 1) * (ny + 1) < (int64_t(1) << 24) && nz + 1 < 65536, FX_EINVAL,
             "transport planes are limited to 2^24 faces and 65534 cells along z (grid %dx%dx%d)", g.nx, g.ny, g.nz);
  return FX_OK;
}

int check_sources(const fx_sources* s) {
  FX_REQUIRE(s, FX_EINVAL, "null source table");
  FX_REQUIRE(s->count >= 0 && s->count <= 8, FX_EINVAL, "At most eight sources are supported");
  for (int i = 0; i < s->count; ++i) {
    FX_REQUIRE(s->table[16 * i + 3] > 0.0f, FX_EINVAL, "source %d radius must be positive", i);
    for (int j = 0; j < 16; ++j)
      FX_REQUIRE(std::isfinite(s->table[16 * i + j]), FX_EINVAL, "source %d has a non-finite entry", i);
  }
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_transport_fast_paths(int32_t enable) {
  return g_fast_paths.exchange(enable ? 1 : 0);
}

int fx_face_predict(fx_grid g, float dt, const float* u, const float* v, const float* w, float* pu, float* pv,
                    float* pw, const uint8_t* solid, const uint8_t* clearance, void* stream) {
  if (int e = check_transport_grid(g)) return e;
  FX_REQUIRE(u && v && w && pu && pv && pw, FX_EINVAL, "null face array");
  FX_REQUIRE(pu != u && pv != v && pw != w && pu != v && pu != w && pv != u && pv != w && pw != u && pw != v,
             FX_EINVAL, "Advection input/output textures must be distinct");
  FX_REQUIRE(std::isfinite(dt), FX_EINVAL, "dt must be finite");
  GridDev G = dev_grid(g);
  Vel V{u, v, w};
  FaceIO F{{u, v, w}, {pu, pv, pw}, {nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  const dim3 grid = site_grid(g.nx + 1, g.ny + 1, g.nz + 1);
  if (solid)
    hipLaunchKernelGGL((k_face_predict<true, false>), grid, kSiteBlock, 0, S(stream), G, V, dt, F, solid, clearance);
  else
    hipLaunchKernelGGL((k_face_predict<false, false>), grid, kSiteBlock, 0, S(stream), G, V, dt, F, solid, clearance);
  FX_LAUNCHED("fx_face_predict");
  return FX_OK;
}

int fx_face_predict_bounds(fx_grid g, float dt, const float* u, const float* v, const float* w, float* const* pred,
                           float* const* lo, float* const* hi, const uint8_t* solid, const uint8_t* clearance,
                           void* stream) {
  if (int e = check_transport_grid(g)) return e;
  FX_REQUIRE(u && v && w && pred && lo && hi, FX_EINVAL, "null face array");
  const float* ins[3] = {u, v, w};
  for (int a = 0; a < 3; ++a) {
    FX_REQUIRE(pred[a] && lo[a] && hi[a], FX_EINVAL, "null face array");
    for (int b = 0; b < 3; ++b)
      FX_REQUIRE(pred[a] != ins[b] && lo[a] != ins[b] && hi[a] != ins[b], FX_EINVAL,
                 "Advection input/output textures must be distinct");
    FX_REQUIRE(pred[a] != lo[a] && pred[a] != hi[a] && lo[a] != hi[a], FX_EINVAL, "bounds arrays must be distinct");
  }
  FX_REQUIRE(std::isfinite(dt), FX_EINVAL, "dt must be finite");
  GridDev G = dev_grid(g);
  Vel V{u, v, w};
  FaceIO F{{u, v, w}, {pred[0], pred[1], pred[2]}, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}};
  const dim3 grid = site_grid(g.nx + 1, g.ny + 1, g.nz + 1);
  if (solid)
    hipLaunchKernelGGL((k_face_predict<true, true>), grid, kSiteBlock, 0, S(stream), G, V, dt, F, solid, clearance);
  else
    hipLaunchKernelGGL((k_face_predict<false, true>), grid, kSiteBlock, 0, S(stream), G, V, dt, F, solid, clearance);
  FX_LAUNCHED("fx_face_predict_bounds");
  return FX_OK;
}

int fx_velocity_advect(const fx_velocity_args* a, void* stream) {
  FX_REQUIRE(a, FX_EINVAL, "null arguments");
  if (int e = check_transport_grid(a->g)) return e;
  if (int e = check_sources(a->sources)) return e;
  FX_REQUIRE(a->u && a->v && a->w && a->ou && a->ov && a->ow, FX_EINVAL, "null face array");
  const float* ins[6] = {a->u, a->v, a->w, a->pu, a->pv, a->pw};
  const float* outs[3] = {a->ou, a->ov, a->ow};
  for (int o = 0; o < 3; ++o)
    for (int k = 0; k < 6; ++k)
      FX_REQUIRE(outs[o] != ins[k], FX_EINVAL, "Advection input/output textures must be distinct");
  FX_REQUIRE(!a->maccormack || (a->pu && a->pv && a->pw), FX_EINVAL, "MacCormack needs face predictors");
  FX_REQUIRE(!a->density == !a->temperature, FX_EINVAL, "buoyancy needs both density and temperature");
  GridDev G = dev_grid(a->g);
  Vel V{a->u, a->v, a->w};
  hipStream_t s = S(stream);
  const fx_grid& g = a->g;
  const bool bounds = a->maccormack && a->lu && a->lv && a->lw && a->hu && a->hv && a->hw;
  VelArgs A{G,
            V,
            a->dt,
            {a->pu, a->pv, a->pw},
            {a->ou, a->ov, a->ow},
            a->density,
            a->temperature,
            a->thermal_lift,
            a->density_weight,
            a->solid,
            a->solid ? a->clearance : nullptr,
            {bounds ? a->lu : nullptr, bounds ? a->lv : nullptr, bounds ? a->lw : nullptr},
            {bounds ? a->hu : nullptr, bounds ? a->hv : nullptr, bounds ? a->hw : nullptr}};
  const dim3 grid = site_grid(g.nx + 1, g.ny + 1, g.nz + 1);
#define FX_VEL(SO, MC, BO) hipLaunchKernelGGL((k_velocity<SO, MC, BO>), grid, kSiteBlock, 0, s, A, *a->sources)
  if (a->solid) {
    if (bounds) FX_VEL(true, true, true);
    else if (a->maccormack) FX_VEL(true, true, false);
    else FX_VEL(true, false, false);
  } else {
    if (bounds) FX_VEL(false, true, true);
    else if (a->maccormack) FX_VEL(false, true, false);
    else FX_VEL(false, false, false);
  }
#undef FX_VEL
  FX_LAUNCHED("fx_velocity_advect");
  return FX_OK;
}

int fx_scalar_predict(fx_grid g, float dt, const float* u, const float* v, const float* w, int32_t channels,
                      const float* const* fields, float* const* preds, const uint8_t* solid,
                      const uint8_t* clearance, void* stream) {
  if (int e = check_transport_grid(g)) return e;
  FX_REQUIRE(channels >= 1 && channels <= 3, FX_EINVAL, "channels must be 1..3");
  FX_REQUIRE(u && v && w && fields && preds, FX_EINVAL, "null array");
  ScalarSet set{};
  for (int k = 0; k < channels; ++k) {
    FX_REQUIRE(fields[k] && preds[k] && fields[k] != preds[k], FX_EINVAL,
               "Advection input/output textures must be distinct");
    set.f[k] = fields[k];
    set.o[k] = preds[k];
  }
  dim3 gr = site_grid(g.nx, g.ny, g.nz);
  const dim3 kBlock = kSiteBlock;
  Vel V{u, v, w};
  if (solid)
    hipLaunchKernelGGL((k_scalar_predict<true>), gr, kBlock, 0, S(stream), dev_grid(g), V, dt, channels, set, solid,
                       clearance);
  else
    hipLaunchKernelGGL((k_scalar_predict<false>), gr, kBlock, 0, S(stream), dev_grid(g), V, dt, channels, set, solid,
                       (const uint8_t*)nullptr);
  FX_LAUNCHED("fx_scalar_predict");
  return FX_OK;
}

}  // extern "C"

namespace fx {

static int scalar_predict_bounds(fx_grid g, float dt, const float* u, const float* v, const float* w,
                                 int32_t channels, const float* const* fields, float* const* preds,
                                 float* const* lo, float* const* hi, const uint8_t* solid, const uint8_t* clearance,
                                 const uint8_t* bricks, void* stream) {
  if (int e = check_transport_grid(g)) return e;
  FX_REQUIRE(channels >= 1 && channels <= 3, FX_EINVAL, "channels must be 1..3");
  FX_REQUIRE(u && v && w && fields && preds && lo && hi, FX_EINVAL, "null array");
  ScalarSet set{};
  for (int k = 0; k < channels; ++k) {
    FX_REQUIRE(fields[k] && preds[k] && lo[k] && hi[k], FX_EINVAL, "null channel %d", k);
    for (int j = 0; j < channels; ++j)
      FX_REQUIRE(preds[k] != fields[j] && lo[k] != fields[j] && hi[k] != fields[j], FX_EINVAL,
                 "Advection input/output textures must be distinct");
    FX_REQUIRE(preds[k] != lo[k] && preds[k] != hi[k] && lo[k] != hi[k], FX_EINVAL, "bounds arrays must be distinct");
    set.f[k] = fields[k];
    set.o[k] = preds[k];
    set.lo[k] = lo[k];
    set.hi[k] = hi[k];
  }
  set.bricks = BrickMask{bricks, active_bricks_along(g.nx), active_bricks_along(g.ny)};
  dim3 gr = site_grid(g.nx, g.ny, g.nz);
  const dim3 kBlock = kSiteBlock;
  Vel V{u, v, w};
  if (solid)
    hipLaunchKernelGGL((k_scalar_predict<true, true>), gr, kBlock, 0, S(stream), dev_grid(g), V, dt, channels, set,
                       solid, clearance);
  else
    hipLaunchKernelGGL((k_scalar_predict<false, true>), gr, kBlock, 0, S(stream), dev_grid(g), V, dt, channels, set,
                       solid, (const uint8_t*)nullptr);
  FX_LAUNCHED("fx_scalar_predict_bounds");
  return FX_OK;
}

static int scalar_advect(const fx_scalar_args* a, const uint8_t* bricks, uint8_t* const* live, void* stream) {
  FX_REQUIRE(a, FX_EINVAL, "null arguments");
  if (int e = check_transport_grid(a->g)) return e;
  if (int e = check_sources(a->sources)) return e;
  FX_REQUIRE(a->channels >= 1 && a->channels <= 3, FX_EINVAL, "channels must be 1..3");
  FX_REQUIRE(a->u && a->v && a->w, FX_EINVAL, "null velocity");
  ScalarArgs A{};
  A.g = dev_grid(a->g);
  A.V = Vel{a->u, a->v, a->w};
  A.dt = a->dt;
  A.channels = a->channels;
  A.solid = a->solid;
  A.clear = a->solid ? a->clearance : nullptr;
  A.zero_ext = a->boundary;
  FX_REQUIRE(a->boundary == 0 || a->boundary == 1, FX_EINVAL, "boundary must be 0 (closed) or 1 (zero)");
  FX_REQUIRE(!(a->boundary == 1 && (a->maccormack || a->solid)), FX_EINVAL,
             "zero-extension transport is semi-Lagrangian without solids");
  for (int k = 0; k < a->channels; ++k) {
    FX_REQUIRE(a->field[k] && a->out[k], FX_EINVAL, "null channel %d", k);
    FX_REQUIRE(!a->maccormack || a->pred[k], FX_EINVAL, "MacCormack needs predictors");
    FX_REQUIRE(a->channel_id[k] >= 0 && a->channel_id[k] <= 2, FX_EINVAL, "channel id must be 0..2");
    for (int j = 0; j < a->channels; ++j)
      FX_REQUIRE(a->out[k] != a->field[j] && a->out[k] != (a->maccormack ? a->pred[j] : nullptr), FX_EINVAL,
                 "Advection input/output textures must be distinct");
    A.f[k] = a->field[k];
    A.pred[k] = a->pred[k];
    A.out[k] = a->out[k];
    A.keep[k] = expf(-a->decay[k] * a->dt);
    A.chan[k] = a->channel_id[k];
    A.live[k] = live ? live[k] : nullptr;
  }
  A.bricks = BrickMask{bricks, active_bricks_along(a->g.nx), active_bricks_along(a->g.ny)};
  bool bounds = a->maccormack != 0 && a->boundary == 0;
  for (int k = 0; k < a->channels; ++k) bounds = bounds && a->lo[k] && a->hi[k];
  for (int k = 0; bounds && k < a->channels; ++k) {
    A.lo[k] = a->lo[k];
    A.hi[k] = a->hi[k];
    for (int j = 0; j < a->channels; ++j)
      FX_REQUIRE(a->out[j] != a->lo[k] && a->out[j] != a->hi[k], FX_EINVAL,
                 "Advection input/output textures must be distinct");
  }
  dim3 gr = site_grid(a->g.nx, a->g.ny, a->g.nz);
  const dim3 kBlock = kSiteBlock;
  hipStream_t s = S(stream);
  if (a->solid) {
    if (bounds) hipLaunchKernelGGL((k_scalar_advect<true, true, true>), gr, kBlock, 0, s, A, *a->sources);
    else if (a->maccormack) hipLaunchKernelGGL((k_scalar_advect<true, true>), gr, kBlock, 0, s, A, *a->sources);
    else hipLaunchKernelGGL((k_scalar_advect<true, false>), gr, kBlock, 0, s, A, *a->sources);
  } else {
    if (bounds) hipLaunchKernelGGL((k_scalar_advect<false, true, true>), gr, kBlock, 0, s, A, *a->sources);
    else if (a->maccormack) hipLaunchKernelGGL((k_scalar_advect<false, true>), gr, kBlock, 0, s, A, *a->sources);
    else hipLaunchKernelGGL((k_scalar_advect<false, false>), gr, kBlock, 0, s, A, *a->sources);
  }
  FX_LAUNCHED("fx_scalar_advect");
  return FX_OK;
}

}  // namespace fx

extern "C" {

int fx_scalar_predict_bounds(fx_grid g, float dt, const float* u, const float* v, const float* w, int32_t channels,
                             const float* const* fields, float* const* preds, float* const* lo, float* const* hi,
                             const uint8_t* solid, const uint8_t* clearance, void* stream) {
  return scalar_predict_bounds(g, dt, u, v, w, channels, fields, preds, lo, hi, solid, clearance, nullptr, stream);
}

int fx_scalar_predict_bounds_active(fx_grid g, float dt, const float* u, const float* v, const float* w,
                                    int32_t channels, const float* const* fields, float* const* preds,
                                    float* const* lo, float* const* hi, const uint8_t* solid,
                                    const uint8_t* clearance, const uint8_t* bricks, void* stream) {
  FX_REQUIRE(bricks, FX_EINVAL, "active predictor: null brick mask");
  return scalar_predict_bounds(g, dt, u, v, w, channels, fields, preds, lo, hi, solid, clearance, bricks, stream);
}

int fx_scalar_advect(const fx_scalar_args* a, void* stream) { return scalar_advect(a, nullptr, nullptr, stream); }

int fx_scalar_advect_active(const fx_scalar_args* a, const uint8_t* bricks, uint8_t* const* live, void* stream) {
  FX_REQUIRE(bricks && live, FX_EINVAL, "active transport: null brick mask or live masks");
  FX_REQUIRE(a, FX_EINVAL, "null arguments");
  FX_REQUIRE(a->channels >= 1 && a->channels <= 3, FX_EINVAL, "channels must be 1..3");
  for (int k = 0; k < a->channels; ++k) FX_REQUIRE(live[k], FX_EINVAL, "active transport: null live mask %d", k);
  return scalar_advect(a, bricks, live, stream);
}

int fx_combustion(int64_t n, float dt, float ignition, float burn_rate, float heat_yield, float smoke_yield,
                  const float* fuel, const float* temperature, const float* density, float* fuel_out,
                  float* temperature_out, float* density_out, float* flame, void* stream) {
  FX_REQUIRE(n >= 0 && fuel && temperature && density && fuel_out && temperature_out && density_out && flame,
             FX_EINVAL, "combustion: null array");
  FX_REQUIRE(dt > 0.0f, FX_EINVAL, "dt must be positive");
  FX_REQUIRE(fuel_out != fuel && temperature_out != temperature && density_out != density, FX_EINVAL,
             "Advection input/output textures must be distinct");
  if (n == 0) return FX_OK;
  hipLaunchKernelGGL(k_combustion, dim3((n + 255) / 256), dim3(256), 0, S(stream), n, dt, ignition, burn_rate,
                     heat_yield, smoke_yield, fuel, temperature, density, fuel_out, temperature_out, density_out,
                     flame);
  FX_LAUNCHED("fx_combustion");
  return FX_OK;
}

int fx_combustion_active(fx_grid g, float dt, float ignition, float burn_rate, float heat_yield, float smoke_yield,
                         const float* fuel, const float* temperature, const float* density, float* fuel_out,
                         float* temperature_out, float* density_out, float* flame, const uint8_t* bricks,
                         uint8_t* const* live, void* stream) {
  if (int e = check_transport_grid(g)) return e;
  FX_REQUIRE(fuel && temperature && density && fuel_out && temperature_out && density_out && flame, FX_EINVAL,
             "combustion: null array");
  FX_REQUIRE(bricks && live && live[0] && live[1] && live[2] && live[3], FX_EINVAL,
             "active combustion: null brick mask or live masks");
  FX_REQUIRE(dt > 0.0f, FX_EINVAL, "dt must be positive");
  FX_REQUIRE(fuel_out != fuel && temperature_out != temperature && density_out != density, FX_EINVAL,
             "Advection input/output textures must be distinct");
  const int n = g.nx * g.ny * g.nz;
  const BrickMask m{bricks, active_bricks_along(g.nx), active_bricks_along(g.ny)};
  hipLaunchKernelGGL(k_combustion_active, dim3((n + 255) / 256), dim3(256), 0, S(stream), g.nx, g.ny, n, dt,
                     ignition, burn_rate, heat_yield, smoke_yield, fuel, temperature, density, fuel_out,
                     temperature_out, density_out, flame, m, live[0], live[1], live[2], live[3]);
  FX_LAUNCHED("fx_combustion_active");
  return FX_OK;
}

int fx_build_clearance(fx_grid g, const uint8_t* solid, uint8_t* clearance, uint8_t* scratch, void* stream) {
  if (int e = check_grid(g)) return e;
  FX_REQUIRE(solid && clearance && scratch && clearance != scratch && clearance != solid && scratch != solid,
             FX_EINVAL, "clearance: needs solid, output and scratch arrays");
  dim3 gr = grid3(g.nx, g.ny, g.nz, kBlock);
  hipStream_t s = S(stream);
  const int r = kClearanceRadius;
  hipLaunchKernelGGL(k_clearance, gr, kBlock, 0, s, g.nx, g.ny, g.nz, 0, solid, scratch, r, 0, 0, 0, g.nx, g.ny);
  hipLaunchKernelGGL(k_clearance, gr, kBlock, 0, s, g.nx, g.ny, g.nz, 1, (const uint8_t*)scratch, clearance, r, 0, 0,
                     0, g.nx, g.ny);
  hipLaunchKernelGGL(k_clearance, gr, kBlock, 0, s, g.nx, g.ny, g.nz, 2, (const uint8_t*)clearance, scratch, r, 0, 0,
                     0, g.nx, g.ny);
  FX_HIP(hipMemcpyAsync(clearance, scratch, size_t(g.nx) * g.ny * g.nz, hipMemcpyDeviceToDevice, s));
  FX_LAUNCHED("fx_build_clearance");
  return FX_OK;
}

int fx_max_abs(int32_t count, const float* const* arrays, const int64_t* sizes, float* out, void* stream) {
  FX_REQUIRE(count >= 1 && count <= 8 && arrays && sizes && out, FX_EINVAL, "max_abs: 1..8 arrays");
  MaxArgs A{};
  for (int k = 0; k < count; ++k) {
    FX_REQUIRE(arrays[k] && sizes[k] >= 0, FX_EINVAL, "max_abs: bad array %d", k);
    A.a[k] = arrays[k];
    A.n[k] = sizes[k];
  }
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_max_init, dim3(1), dim3(64), 0, s, out, count);
  hipLaunchKernelGGL(k_max_abs, dim3(1024, count), dim3(256), 0, s, A, out);
  FX_LAUNCHED("fx_max_abs");
  return FX_OK;
}

}  // extern "C"
