// Learned pressure preconditioner: a 3-level 3D U-Net (16/32/64 channels,
// 3x3x3 convolutions, ReLU) on bf16 activations with fp32 accumulation on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).  No reference
// counterpart: the layer list is the one of the task's learned-preconditioner
// configuration (2 inputs: divergence and solid mask, 1 output: pressure).
//
// Layout: activations are channels-last bf16, act[v * C + c] with the voxel
// index v = x + nx*(y + ny*z).  Convolution weights are packed per output
// channel, wt[co * kp + k] with k = tap*cin + ci, tap = (dz+1)*9 + (dy+1)*3 +
// (dx+1), kp = 27*cin rounded up to 32 (zero tail).
//
// Forward convolution = implicit GEMM, M = voxels, N = cout, K = 27*cin.  A
// wave owns 2 tiles of 16 consecutive voxels x all output channels; per
// 32-wide K step every lane loads the 8 channels of one (voxel, tap) as one
// 16-byte vector (A operand) and 8 weights of one output channel (B operand).
// The input may be the concatenation of a nearest-upsampled coarse tensor and
// a skip tensor (the decoder's "up + concat"), resolved in the A load, so
// neither is materialised.  The data gradient is the same kernel with the
// weights flipped in space and transposed in channels; the weight gradient
// is a GEMM over voxels (K = voxels) on channel-major copies of its operands.
#include <algorithm>

#include "common.hpp"

namespace fx {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4_t;

struct ConvDev {
  int nx, ny, nz;
  int c0, c1, cin, cout, kp;
  int up0, relu;
};

// round to nearest even (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ uint16_t bf16_bits(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
__device__ __forceinline__ float bf16_val(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }

// 8 channels [ci0, ci0+8) of voxel (x, y, z) of the concatenated input; zero
// outside the grid (zero padding)
__device__ __forceinline__ uint4 conv_load_a(const ConvDev& d, const uint16_t* __restrict__ in0,
                                             const uint16_t* __restrict__ in1, int x, int y, int z, int ci0) {
  if (x < 0 || y < 0 || z < 0 || x >= d.nx || y >= d.ny || z >= d.nz) return make_uint4(0, 0, 0, 0);
  if (ci0 < d.c0) {
    int64_t v;
    if (d.up0) {
      const int hx = d.nx >> 1, hy = d.ny >> 1;
      v = (int64_t(z >> 1) * hy + (y >> 1)) * hx + (x >> 1);
    } else {
      v = (int64_t(z) * d.ny + y) * d.nx + x;
    }
    return *reinterpret_cast<const uint4*>(in0 + v * d.c0 + ci0);
  }
  const int64_t v = (int64_t(z) * d.ny + y) * d.nx + x;
  return *reinterpret_cast<const uint4*>(in1 + v * d.c1 + (ci0 - d.c0));
}

__device__ __forceinline__ bf16x8_t as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8_t, u); }

// Forward convolution, LDS-tiled: a block owns 16 (x) x 4 (y) x 4 (z) output
// voxels; wave w owns plane w of the tile as four 16-voxel M tiles (one per
// row) times all output channels.  Per input-channel chunk (8, 16 or 32
// channels) the block stages the 18 x 6 x 6 halo tile in LDS (zero outside
// the grid; upsampling and concatenation resolved while staging) together
// with the chunk's weights in B-operand order.
//
// The K loop runs over "sets": groups of 32/cw (dz, dx) tap offsets (lane
// group lg takes offset (8 lg)/cw of the set and channels (8 lg) mod cw).  For
// a set a wave reads the six A rows y = -1 .. 4 of its plane once and feeds
// them to the three dy taps (output row t with tap dy uses row t + dy), so a
// set costs 6 A reads + 3 B reads per 12 x NT MFMAs.  With one 16-channel
// output tile (FULLB) the weights of all 27 taps of the chunk sit in LDS and
// a set may mix dz planes; with more output tiles LDS holds one dz slice at a
// time and sets group dx offsets only.  A voxel occupies 16 bytes more than
// its channels in LDS, which spreads the 16 lanes of a fragment read over all
// 64 banks.
constexpr int kTX = 16, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;  // 648 voxels
constexpr int kChunk = 32;              // input channels per staged chunk
constexpr int kVoxBytes = kChunk * 2 + 16;
constexpr int kFrag = 1024;             // one B operand: 64 lanes x 16 bytes

// LDS bytes of the staged weights: FULLB 9 sets (cw = 32) x 3 dy, else 3 sets x 3 dy, NT tiles each
template <int NT, bool FULLB>
constexpr int conv_b_bytes() {
  return (FULLB ? 27 : 9) * NT * kFrag;
}

// CIN > 0: input channel count known at compile time (chunk loops unrolled,
// LDS offsets immediate); CIN = 0: any multiple of 8 (d.cin)
template <int CIN, int NT>
__global__ __launch_bounds__(256) void k_conv_fwd(ConvDev d, const uint16_t* __restrict__ in0,
                                                  const uint16_t* __restrict__ in1, const uint16_t* __restrict__ wt,
                                                  const float* __restrict__ bias, uint16_t* __restrict__ out,
                                                  int tiles_x, int tiles_y) {
  constexpr bool FULLB = NT == 1;
  __shared__ __attribute__((aligned(16))) unsigned char tile[kHalo * kVoxBytes];
  __shared__ __attribute__((aligned(16))) unsigned char wlds[conv_b_bytes<NT, FULLB>()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int b = blockIdx.x;
  const int tx = b % tiles_x, rest = b / tiles_x, ty = rest % tiles_y, tz = rest / tiles_y;
  const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;
  const int cin = CIN > 0 ? CIN : d.cin;
  floatx4_t acc[kTY][NT];
#pragma unroll
  for (int t = 0; t < kTY; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[t][n] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
  // element offsets of this thread's halo voxels (one per 256 threads) in in0
  // (upsampled when up0) and in1; -1: outside the grid (zero padding)
  constexpr int kSlotsA = (kHalo + 255) / 256;
  int64_t v0[kSlotsA], v1[kSlotsA];
#pragma unroll
  for (int q = 0; q < kSlotsA; ++q) {
    const int hv = int(threadIdx.x) + 256 * q;
    const int hx = hv % kHX, hr = hv / kHX, hy = hr % kHY, hz = hr / kHY;
    const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = z0 - 1 + hz;
    v0[q] = v1[q] = -1;
    if (hv < kHalo && x >= 0 && y >= 0 && z >= 0 && x < d.nx && y < d.ny && z < d.nz) {
      const int64_t v = (int64_t(z) * d.ny + y) * d.nx + x;
      v1[q] = v * d.c1;
      v0[q] = (d.up0 ? (int64_t(z >> 1) * (d.ny >> 1) + (y >> 1)) * (d.nx >> 1) + (x >> 1) : v) * d.c0;
    }
  }
#pragma unroll
  for (int cbase = 0; cbase < (CIN > 0 ? CIN : 4096); cbase += kChunk) {
    if (CIN == 0 && cbase >= cin) break;
    const int cw = min(kChunk, cin - cbase);  // 8, 16 or 32
    const int groups = cw >> 3;
    const int shift = cw == 32 ? 5 : (cw == 16 ? 4 : 3);
    const int per_set = 32 >> shift;            // tap offsets per set
    const int kof = (lg * 8) >> shift;          // this lane group's offset within a set
    const int cg = (lg * 8) & (cw - 1);         // and its channel offset
    const int noff = FULLB ? 9 : 3;             // (dz, dx) offsets per pass
    const int nsets = (noff + per_set - 1) / per_set;
    // staging: every load of a fill is issued before the first LDS store, so a
    // fill costs one memory latency (loads in registers, then stores)
    uint4 va[kSlotsA][4];
#pragma unroll
    for (int q = 0; q < kSlotsA; ++q)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        va[q][g] = make_uint4(0, 0, 0, 0);
        const int ci = cbase + 8 * g;
        if (g < groups) {
          if (ci < d.c0) {
            if (v0[q] >= 0) va[q][g] = *reinterpret_cast<const uint4*>(in0 + v0[q] + ci);
          } else if (v1[q] >= 0) {
            va[q][g] = *reinterpret_cast<const uint4*>(in1 + v1[q] + (ci - d.c0));
          }
        }
      }
    __syncthreads();  // previous chunk fully consumed
#pragma unroll
    for (int q = 0; q < kSlotsA; ++q) {
      const int hv = int(threadIdx.x) + 256 * q;
      if (hv < kHalo)
#pragma unroll
        for (int g = 0; g < 4; ++g)
          if (g < groups) *reinterpret_cast<uint4*>(tile + hv * kVoxBytes + g * 16) = va[q][g];
    }
    for (int pass = 0; pass < (FULLB ? 1 : 3); ++pass) {
      // weights of this pass in B-operand order: fragment ((set * 3 + dy) * NT + n),
      // lane-major; item (fragment, channel li) gathers the four lane groups
      constexpr int kItems = conv_b_bytes<NT, FULLB>() / kFrag * 16;
      constexpr int kSlotsB = (kItems + 255) / 256;
      const int nitems = nsets * 3 * NT * 16;
      uint4 vb[kSlotsB][4];
#pragma unroll
      for (int q = 0; q < kSlotsB; ++q) {
        const int item = int(threadIdx.x) + 256 * q;
        const int frag = item >> 4, co_lo = item & 15;
        const int n = frag % NT, sd = frag / NT, dyi = sd % 3, set = sd / 3;
        const uint16_t* wrow = wt + int64_t(16 * n + co_lo) * d.kp + cbase + dyi * 3 * cin;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          vb[q][g] = make_uint4(0, 0, 0, 0);
          const int k = set * per_set + ((g * 8) >> shift);
          if (item < nitems && k < noff) {
            const int dzi = FULLB ? k / 3 : pass, dxi = FULLB ? k - 3 * (k / 3) : k;
            vb[q][g] = *reinterpret_cast<const uint4*>(wrow + (dzi * 9 + dxi) * cin + ((g * 8) & (cw - 1)));
          }
        }
      }
      if (!FULLB && pass > 0) __syncthreads();  // previous dz slice consumed
#pragma unroll
      for (int q = 0; q < kSlotsB; ++q) {
        const int item = int(threadIdx.x) + 256 * q;
        if (item < nitems)
#pragma unroll
          for (int g = 0; g < 4; ++g)
            *reinterpret_cast<uint4*>(wlds + (item >> 4) * kFrag + (g * 16 + (item & 15)) * 16) = vb[q][g];
      }
      __syncthreads();
#pragma unroll
      for (int set = 0; set < (FULLB ? 9 : 3); ++set) {
        if (set >= nsets) break;
        // this lane group's tap offset (padding offsets read offset 0: their weights are zero)
        int k = set * per_set + kof;
        if (k >= noff) k = 0;
        const int dz = FULLB ? k / 3 - 1 : pass - 1, dx = FULLB ? k % 3 - 1 : k - 1;
        const int base = (((wave + 1 + dz) * kHY) * kHX + li + 1 + dx) * kVoxBytes + cg * 2;
        uint4 a[kTY + 2];
#pragma unroll
        for (int r = 0; r < kTY + 2; ++r) a[r] = *reinterpret_cast<const uint4*>(tile + base + r * kHX * kVoxBytes);
#pragma unroll
        for (int dyi = 0; dyi < 3; ++dyi) {
          uint4 bw[NT];
#pragma unroll
          for (int n = 0; n < NT; ++n)
            bw[n] = *reinterpret_cast<const uint4*>(wlds + ((set * 3 + dyi) * NT + n) * kFrag + lane * 16);
#pragma unroll
          for (int t = 0; t < kTY; ++t)
#pragma unroll
            for (int n = 0; n < NT; ++n)
              acc[t][n] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a[t + dyi]), as_bf16x8(bw[n]), acc[t][n], 0, 0, 0);
        }
      }
    }
  }
  // lane (li, lg) holds x = x0 + 4*lg + j of row t, output channel 16*n + li
  const int z = z0 + wave;
  if (z >= d.nz) return;
#pragma unroll
  for (int t = 0; t < kTY; ++t) {
    const int y = y0 + t;
    if (y >= d.ny) continue;
    uint16_t* orow = out + ((int64_t(z) * d.ny + y) * d.nx + x0 + 4 * lg) * d.cout + li;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const float bv = bias ? bias[16 * n + li] : 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (x0 + 4 * lg + j < d.nx) {
          float r = acc[t][n][j] + bv;
          if (d.relu) r = fmaxf(r, 0.0f);
          orow[j * d.cout + 16 * n] = bf16_bits(r);
        }
      }
    }
  }
}

// Block barrier ordering LDS only: global stores in flight (conv outputs,
// never read back by the block) are not waited for, unlike __syncthreads().
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Pipelined form of k_conv_fwd for a compile-time channel count: a block
// owns `tiles` consecutive 16 x 4 x 4 tiles along z of one (x, y) column and
// walks the steps (tile, chunk, dz pass).  While a step computes from LDS, the
// global loads of the next step's halo (and weights) are already in flight in
// registers.  With one output tile and one chunk (BREG) every wave keeps the
// B operands of all its K steps in registers for the whole block and LDS
// holds only the halo tile.  Same sums per output as k_conv_fwd.
template <int CIN, int NT>
__global__ __launch_bounds__(256) void k_conv_pipe(ConvDev d, const uint16_t* __restrict__ in0,
                                                   const uint16_t* __restrict__ in1,
                                                   const uint16_t* __restrict__ wt, const float* __restrict__ bias,
                                                   uint16_t* __restrict__ out, int tiles_x, int tiles_y, int tiles) {
  constexpr bool FULLB = NT == 1;
  constexpr int NCH = (CIN + 31) / 32, PASSES = FULLB ? 1 : 3;
  constexpr bool BREG = FULLB && NCH == 1;
  constexpr int kSlotsA = (kHalo + 255) / 256;
  constexpr int kItems = conv_b_bytes<NT, FULLB>() / kFrag * 16, kSlotsB = BREG ? 1 : (kItems + 255) / 256;
  constexpr int COUT = 16 * NT;
  __shared__ __attribute__((aligned(16))) unsigned char tile[kHalo * kVoxBytes];
  __shared__ __attribute__((aligned(16))) unsigned char wlds[BREG ? 16 : conv_b_bytes<NT, FULLB>()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int b = blockIdx.x;
  const int tx = b % tiles_x, rest = b / tiles_x, ty = rest % tiles_y, tzg = rest / tiles_y;
  const int x0 = tx * kTX, y0 = ty * kTY;
  const int nsteps = tiles * NCH * PASSES;
  auto chunk_w = [](int c) { return CIN % 32 == 0 ? 32 : (c == NCH - 1 ? CIN % 32 : 32); };
  auto tile_z0 = [&](int k) { return (tzg * tiles + k) * kTZ; };
  floatx4_t acc[kTY][NT];
#pragma unroll
  for (int t = 0; t < kTY; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[t][n] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
  // this thread's halo voxels: element offsets at tile 0 and per-tile voxel
  // increments (tile k lies k * 4 planes further along z; 2 coarse planes when
  // in0 is upsampled)
  const int64_t plane = int64_t(d.nx) * d.ny;
  const int64_t step0 = d.up0 ? 2 * int64_t(d.nx >> 1) * (d.ny >> 1) : 4 * plane;
  int64_t hv0[kSlotsA], hv1[kSlotsA];
  int hzs[kSlotsA];
  bool hxy[kSlotsA];
#pragma unroll
  for (int q = 0; q < kSlotsA; ++q) {
    const int hv = int(threadIdx.x) + 256 * q;
    const int hx = hv % kHX, hr = hv / kHX, hy = hr % kHY, hz = hr / kHY;
    const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = tile_z0(0) - 1 + hz;
    hxy[q] = hv < kHalo && x >= 0 && y >= 0 && x < d.nx && y < d.ny;
    hzs[q] = z;
    const int xc = max(x, 0), yc = max(y, 0);  // outside voxels are never loaded
    const int64_t v = (int64_t(z) * d.ny + yc) * d.nx + xc;
    hv1[q] = v * d.c1;
    hv0[q] = (d.up0 ? (int64_t(z >> 1) * (d.ny >> 1) + (yc >> 1)) * (d.nx >> 1) + (xc >> 1) : v) * d.c0;
  }
  uint4 va[kSlotsA][4], vb[kSlotsB][4];
  auto load_a = [&](int k, int c) {
    const int cbase = 32 * c, groups = chunk_w(c) >> 3;
    const int kz = k * kTZ;
#pragma unroll
    for (int q = 0; q < kSlotsA; ++q) {
      const int z = hzs[q] + kz;
      const bool inside = hxy[q] && z >= 0 && z < d.nz;
      const int64_t v0 = hv0[q] + k * step0 * d.c0, v1 = hv1[q] + k * 4 * plane * d.c1;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        va[q][g] = make_uint4(0, 0, 0, 0);
        const int ci = cbase + 8 * g;
        if (inside && g < groups)
          va[q][g] = ci < d.c0 ? *reinterpret_cast<const uint4*>(in0 + v0 + ci)
                               : *reinterpret_cast<const uint4*>(in1 + v1 + (ci - d.c0));
      }
    }
  };
  auto store_a = [&](int c) {
    const int groups = chunk_w(c) >> 3;
#pragma unroll
    for (int q = 0; q < kSlotsA; ++q) {
      const int hv = int(threadIdx.x) + 256 * q;
      if (hv < kHalo)
#pragma unroll
        for (int g = 0; g < 4; ++g)
          if (g < groups) *reinterpret_cast<uint4*>(tile + hv * kVoxBytes + g * 16) = va[q][g];
    }
  };
  auto load_b = [&](int c, int pass) {
    const int cbase = 32 * c, cw = chunk_w(c), shift = cw == 32 ? 5 : (cw == 16 ? 4 : 3);
    const int per_set = 32 >> shift, noff = FULLB ? 9 : 3, nsets = (noff + per_set - 1) / per_set;
    const int nitems = nsets * 3 * NT * 16;
#pragma unroll
    for (int q = 0; q < kSlotsB; ++q) {
      const int item = int(threadIdx.x) + 256 * q;
      const int frag = item >> 4, co_lo = item & 15;
      const int n = frag % NT, sd = frag / NT, dyi = sd % 3, set = sd / 3;
      const uint16_t* wrow = wt + int64_t(16 * n + co_lo) * d.kp + cbase + dyi * 3 * CIN;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        vb[q][g] = make_uint4(0, 0, 0, 0);
        const int k = set * per_set + ((g * 8) >> shift);
        if (item < nitems && k < noff) {
          const int dzi = FULLB ? k / 3 : pass, dxi = FULLB ? k - 3 * (k / 3) : k;
          vb[q][g] = *reinterpret_cast<const uint4*>(wrow + (dzi * 9 + dxi) * CIN + ((g * 8) & (cw - 1)));
        }
      }
    }
  };
  auto store_b = [&](int c) {
    const int cw = chunk_w(c), shift = cw == 32 ? 5 : (cw == 16 ? 4 : 3);
    const int per_set = 32 >> shift, noff = FULLB ? 9 : 3, nsets = (noff + per_set - 1) / per_set;
    const int nitems = nsets * 3 * NT * 16;
#pragma unroll
    for (int q = 0; q < kSlotsB; ++q) {
      const int item = int(threadIdx.x) + 256 * q;
      if (item < nitems)
#pragma unroll
        for (int g = 0; g < 4; ++g)
          *reinterpret_cast<uint4*>(wlds + (item >> 4) * kFrag + (g * 16 + (item & 15)) * 16) = vb[q][g];
    }
  };
  // BREG: the wave's B fragments (set, dy) for all K steps, straight from global
  constexpr int kCW0 = CIN >= 32 ? 32 : CIN;
  constexpr int kShift0 = kCW0 == 32 ? 5 : (kCW0 == 16 ? 4 : 3);
  constexpr int kSets0 = (9 + (32 >> kShift0) - 1) / (32 >> kShift0);
  uint4 breg[BREG ? kSets0 * 3 : 1];
  if (BREG) {
#pragma unroll
    for (int set = 0; set < kSets0; ++set)
#pragma unroll
      for (int dyi = 0; dyi < 3; ++dyi) {
        const int k = set * (32 >> kShift0) + ((lg * 8) >> kShift0);
        breg[set * 3 + dyi] = make_uint4(0, 0, 0, 0);
        if (k < 9)
          breg[set * 3 + dyi] = *reinterpret_cast<const uint4*>(
              wt + int64_t(li) * d.kp + ((k / 3) * 9 + dyi * 3 + k % 3) * CIN + ((lg * 8) & (kCW0 - 1)));
      }
  }
  // weights change at every step of the slice form; with FULLB at every chunk
  // change (every step when NCH > 1, else only at the first)
  auto new_b = [&](int s) { return !BREG && (!FULLB || NCH > 1 || s == 0); };
  load_a(0, 0);
  if (new_b(0)) load_b(0, 0);
  for (int s = 0; s < nsteps; ++s) {
    const int pass = s % PASSES, c = (s / PASSES) % NCH, k = s / (PASSES * NCH);
    if (tile_z0(k) >= d.nz) break;  // uniform over the block
    lds_barrier();  // previous step fully consumed
    if (pass == 0) store_a(c);
    if (new_b(s)) store_b(c);
    lds_barrier();
    if (s + 1 < nsteps) {
      const int p1 = (s + 1) % PASSES, c1 = ((s + 1) / PASSES) % NCH, k1 = (s + 1) / (PASSES * NCH);
      if (tile_z0(k1) < d.nz) {
        if (p1 == 0) load_a(k1, c1);
        if (new_b(s + 1)) load_b(c1, p1);
      }
    }
    {
      const int cw = chunk_w(c), shift = cw == 32 ? 5 : (cw == 16 ? 4 : 3);
      const int per_set = 32 >> shift, noff = FULLB ? 9 : 3, nsets = (noff + per_set - 1) / per_set;
      const int kof = (lg * 8) >> shift, cg = (lg * 8) & (cw - 1);
#pragma unroll
      for (int set = 0; set < (FULLB ? 9 : 3); ++set) {
        if (set >= nsets) break;
        int kk = set * per_set + kof;
        if (kk >= noff) kk = 0;  // padding offsets: their weights are zero
        const int dz = FULLB ? kk / 3 - 1 : pass - 1, dx = FULLB ? kk % 3 - 1 : kk - 1;
        const int base = (((wave + 1 + dz) * kHY) * kHX + li + 1 + dx) * kVoxBytes + cg * 2;
        uint4 a[kTY + 2];
#pragma unroll
        for (int r = 0; r < kTY + 2; ++r) a[r] = *reinterpret_cast<const uint4*>(tile + base + r * kHX * kVoxBytes);
#pragma unroll
        for (int dyi = 0; dyi < 3; ++dyi) {
          uint4 bw[NT];
#pragma unroll
          for (int n = 0; n < NT; ++n)
            bw[n] = BREG ? breg[BREG ? set * 3 + dyi : 0]
                         : *reinterpret_cast<const uint4*>(wlds + ((set * 3 + dyi) * NT + n) * kFrag + lane * 16);
#pragma unroll
          for (int t = 0; t < kTY; ++t)
#pragma unroll
            for (int n = 0; n < NT; ++n)
              acc[t][n] =
                  __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a[t + dyi]), as_bf16x8(bw[n]), acc[t][n], 0, 0, 0);
        }
      }
    }
    if (pass == PASSES - 1 && c == NCH - 1) {
      // lane (li, lg) holds x = x0 + 4*lg + j of row t, output channel 16*n + li
      const int z = tile_z0(k) + wave;
      float bv[NT];
#pragma unroll
      for (int n = 0; n < NT; ++n) bv[n] = bias ? bias[16 * n + li] : 0.0f;
#pragma unroll
      for (int t = 0; t < kTY; ++t) {
        const int y = y0 + t;
        if (z < d.nz && y < d.ny) {
          uint16_t* orow = out + ((int64_t(z) * d.ny + y) * d.nx + x0 + 4 * lg) * COUT + li;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            if (x0 + 4 * lg + j < d.nx) {
#pragma unroll
              for (int n = 0; n < NT; ++n) {
                float r = acc[t][n][j] + bv[n];
                if (d.relu) r = fmaxf(r, 0.0f);
                orow[j * COUT + 16 * n] = bf16_bits(r);
              }
            }
          }
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[t][n] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
      }
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[tap][ci][co] = sum_v in(v + off(tap), ci) * g(v, co)
// on channel-major operands inT[ci][v], gT[co][v].  K = voxels in groups of 8
// consecutive x (nx % 8 == 0); a K step takes 4 groups (lane group lg -> one
// group).  A wave owns one (dz, dy) pair, one 16-channel input tile, the three
// dx taps (one load of the row words, shifted by one element for dx = +-1)
// and every output tile; partial sums per voxel chunk go to part[chunk][...].
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 ld_words(const uint16_t* p) { return *reinterpret_cast<const uint4*>(p); }

// elements [1..8] of the 16-element window (cur, next)
__device__ __forceinline__ uint4 shift_up(uint4 cur, uint4 next) {
  return make_uint4((cur.x >> 16) | (cur.y << 16), (cur.y >> 16) | (cur.z << 16), (cur.z >> 16) | (cur.w << 16),
                    (cur.w >> 16) | (next.x << 16));
}
// elements [7..14] of the window (prev, cur)
__device__ __forceinline__ uint4 shift_down(uint4 prev, uint4 cur) {
  return make_uint4((prev.w >> 16) | (cur.x << 16), (cur.x >> 16) | (cur.y << 16), (cur.y >> 16) | (cur.z << 16),
                    (cur.z >> 16) | (cur.w << 16));
}

template <int NT>
__global__ __launch_bounds__(256) void k_conv_wgrad(int nx, int ny, int nz, int cin, int cout,
                                                    const uint16_t* __restrict__ inT,
                                                    const uint16_t* __restrict__ gT, float* __restrict__ part,
                                                    int chunks, int groups_per_chunk, int ci_tiles) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int task = blockIdx.x * 4 + wave;
  const int ntask = 9 * ci_tiles * chunks;
  if (task >= ntask) return;
  const int chunk = task % chunks;
  const int rest = task / chunks;
  const int cit = rest % ci_tiles;
  const int dzy = rest / ci_tiles;
  const int dz = dzy / 3 - 1, dy = dzy % 3 - 1;
  const int64_t V = int64_t(nx) * ny * nz;
  const int gx = nx >> 3;
  const int64_t ngroups = V >> 3;
  const int ci = cit * 16 + li;
  const bool ci_ok = ci < cin;
  const uint16_t* arow = inT + int64_t(ci_ok ? ci : 0) * V;
  floatx4_t acc[3][NT];
#pragma unroll
  for (int s = 0; s < 3; ++s)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[s][n] = floatx4_t{0.0f, 0.0f, 0.0f, 0.0f};
  const int64_t g_begin = int64_t(chunk) * groups_per_chunk;
  for (int gs = 0; gs < groups_per_chunk; gs += 4) {
    const int64_t G = g_begin + gs + lg;
    uint4 a3[3] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
    uint4 b[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) b[n] = make_uint4(0, 0, 0, 0);
    if (G < ngroups) {
      const int64_t row = G / gx;
      const int x0 = int(G - row * gx) << 3;
      const int z = int(row / ny), y = int(row - int64_t(z) * ny);
      const int64_t v0 = row * nx + x0;
#pragma unroll
      for (int n = 0; n < NT; ++n) b[n] = ld_words(gT + int64_t(16 * n + li) * V + v0);
      const int ys = y + dy, zs = z + dz;
      if (ci_ok && ys >= 0 && ys < ny && zs >= 0 && zs < nz) {
        const int64_t base = (int64_t(zs) * ny + ys) * nx + x0;
        const uint4 cur = ld_words(arow + base);
        const uint4 prev = x0 > 0 ? ld_words(arow + base - 8) : make_uint4(0, 0, 0, 0);
        const uint4 next = x0 + 8 < nx ? ld_words(arow + base + 8) : make_uint4(0, 0, 0, 0);
        a3[0] = shift_down(prev, cur);  // dx = -1
        a3[1] = cur;                    // dx = 0
        a3[2] = shift_up(cur, next);    // dx = +1
      }
    }
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int n = 0; n < NT; ++n)
        acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a3[s]), as_bf16x8(b[n]), acc[s][n], 0, 0, 0);
  }
  // lane (li, lg): rows (input channels) cit*16 + 4*lg + j, column (output channel) 16*n + li
  float* pc = part + int64_t(chunk) * 27 * cin * cout;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int tap = (dz + 1) * 9 + (dy + 1) * 3 + s;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = cit * 16 + 4 * lg + j;
        if (c < cin) pc[(int64_t(tap) * cin + c) * cout + 16 * n + li] = acc[s][n][j];
      }
    }
  }
}

// dw (+)= sum over chunks of the partials
__global__ void k_sum_chunks(int64_t n, int chunks, const float* __restrict__ part, float* __restrict__ dw,
                             int accumulate) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    float s = accumulate ? dw[i] : 0.0f;
    for (int c = 0; c < chunks; ++c) s += part[int64_t(c) * n + i];
    dw[i] = s;
  }
}

// channel-major copy of the (virtual, concatenated, upsampled) conv input:
// outT[ci][v], one lane per voxel
__global__ void k_transpose_in(ConvDev d, const uint16_t* __restrict__ in0, const uint16_t* __restrict__ in1,
                               uint16_t* __restrict__ outT, int nvox) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  const int row = v / d.nx, x = v - row * d.nx, z = row / d.ny, y = row - z * d.ny;
  for (int c = 0; c < d.cin; c += 8) {
    const uint4 w = conv_load_a(d, in0, in1, x, y, z, c);
    const uint32_t u[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) outT[int64_t(c + j) * nvox + v] = uint16_t((u[j >> 1] >> (16 * (j & 1))) & 0xffffu);
  }
}

// gT[co][v] = g[v][co] * (act[v][co] > 0 when act is given)
__global__ void k_transpose_grad(int nvox, int cout, const uint16_t* __restrict__ g, const uint16_t* __restrict__ act,
                                 uint16_t* __restrict__ gT) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  for (int c = 0; c < cout; ++c) {
    uint16_t gv = g[int64_t(v) * cout + c];
    if (act && !(bf16_val(act[int64_t(v) * cout + c]) > 0.0f)) gv = 0;
    gT[int64_t(c) * nvox + v] = gv;
  }
}

// g[v][c] *= (act[v][c] > 0), in place (ReLU backward at the stored output)
__global__ void k_relu_mask(int64_t n, uint16_t* __restrict__ g, const uint16_t* __restrict__ act) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    if (!(bf16_val(act[i]) > 0.0f)) g[i] = 0;
}

// db[co] (+)= sum_v g[v][co] (* relu mask); one block per output channel
__global__ void k_bias_grad(int nvox, int cout, const uint16_t* __restrict__ g, const uint16_t* __restrict__ act,
                            float* __restrict__ db, int accumulate) {
  const int co = blockIdx.x;
  float s = 0.0f;
  for (int v = threadIdx.x; v < nvox; v += blockDim.x) {
    const int64_t i = int64_t(v) * cout + co;
    if (!act || bf16_val(act[i]) > 0.0f) s += bf16_val(g[i]);
  }
  __shared__ float part[256];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = (accumulate ? db[co] : 0.0f) + part[0];
}

// 2x2x2 mean of a channels-last tensor (fine grid nx, ny, nz)
// thread = (coarse voxel, 8-channel group): eight 16-byte child loads, one
// 16-byte store; per channel the children are summed in (dz, dy, dx) order
__global__ void k_pool(int nx, int ny, int nz, int c, const uint16_t* __restrict__ in, uint16_t* __restrict__ out) {
  const int hx = nx >> 1, hy = ny >> 1, hz = nz >> 1;
  const int groups = c >> 3;
  const int64_t n = int64_t(hx) * hy * hz * groups;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int g = int(i % groups);
    const int64_t v = i / groups;
    const int X = int(v % hx);
    const int64_t row = v / hx;
    const int Y = int(row % hy), Z = int(row / hy);
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const uint4 a = *reinterpret_cast<const uint4*>(
              in + ((int64_t(2 * Z + dz) * ny + 2 * Y + dy) * nx + 2 * X + dx) * c + 8 * g);
          const uint32_t u[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
          for (int k = 0; k < 8; ++k) s[k] += bf16_val(uint16_t((u[k >> 1] >> (16 * (k & 1))) & 0xffffu));
        }
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = uint32_t(bf16_bits(s[2 * k] * 0.125f)) | (uint32_t(bf16_bits(s[2 * k + 1] * 0.125f)) << 16);
    *reinterpret_cast<uint4*>(out + v * c + 8 * g) = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// out[v][ch] = skip[v][ch] + gpool[parent(v)][ch] / 8   (pool backward + skip sum)
__global__ void k_pool_back_add(int nx, int ny, int nz, int c, const uint16_t* __restrict__ gpool,
                                const uint16_t* __restrict__ skip, uint16_t* __restrict__ out) {
  const int64_t n = int64_t(nx) * ny * nz * c;
  const int hx = nx >> 1, hy = ny >> 1;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int ch = int(i % c);
    const int64_t v = i / c;
    const int x = int(v % nx), y = int((v / nx) % ny), z = int(v / (int64_t(nx) * ny));
    const int64_t pv = (int64_t(z >> 1) * hy + (y >> 1)) * hx + (x >> 1);
    float s = bf16_val(gpool[pv * c + ch]) * 0.125f;
    if (skip) s += bf16_val(skip[i]);
    out[i] = bf16_bits(s);
  }
}

// split the gradient of a concatenated conv input [v][c0 + c1] (fine grid):
// gup[coarse][ch] = sum over the 8 children of g[.][ch] (nearest-upsampling adjoint),
// gskip[v][ch] = g[v][c0 + ch]
__global__ void k_concat_back(int nx, int ny, int nz, int c0, int c1, const uint16_t* __restrict__ g,
                              uint16_t* __restrict__ gup, uint16_t* __restrict__ gskip) {
  const int hx = nx >> 1, hy = ny >> 1, hz = nz >> 1;
  const int cin = c0 + c1;
  const int64_t ncoarse = int64_t(hx) * hy * hz * c0;
  const int64_t nskip = int64_t(nx) * ny * nz * c1;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < ncoarse + nskip;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (i < ncoarse) {
      const int ch = int(i % c0);
      const int64_t v = i / c0;
      const int X = int(v % hx), Y = int((v / hx) % hy), Z = int(v / (int64_t(hx) * hy));
      float s = 0.0f;
      for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
          for (int dx = 0; dx < 2; ++dx)
            s += bf16_val(g[((int64_t(2 * Z + dz) * ny + 2 * Y + dy) * nx + 2 * X + dx) * cin + ch]);
      gup[i] = bf16_bits(s);
    } else {
      const int64_t j = i - ncoarse;
      const int ch = int(j % c1);
      const int64_t v = j / c1;
      gskip[j] = g[v * cin + c0 + ch];
    }
  }
}

// packed input: ch0 = r*scale (scale = 1/rms of r over fluid cells, from ws[q]),
// ch1 = solid indicator, ch2..7 = 0
__global__ void k_unet_input(int64_t n, const float* __restrict__ r, const uint8_t* __restrict__ solid,
                             const double* __restrict__ ws, int q, uint16_t* __restrict__ out) {
  const double ss = ws[q], cnt = ws[q + 1];
  const float scale = (ss > 0.0 && cnt > 0.0) ? float(1.0 / sqrt(ss / cnt)) : 0.0f;
  // one voxel per thread: each store instruction writes 1 KB contiguous
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const bool s = solid && solid[i];
    uint4 w;
    w.x = uint32_t(bf16_bits(s ? 0.0f : r[i] * scale)) | (uint32_t(bf16_bits(s ? 1.0f : 0.0f)) << 16);
    w.y = w.z = w.w = 0;
    *reinterpret_cast<uint4*>(out + i * 8) = w;
  }
}

// ws[q] = sum of r^2 over fluid cells, ws[q+1] = fluid count (per-block partials)
__global__ void k_fluid_norm(int64_t n, const float* __restrict__ r, const uint8_t* __restrict__ solid, double* ws,
                             int q) {
  double s = 0.0, c = 0.0;
  // four voxels per thread and iteration (n % 4 == 0), added in index order
  for (int64_t i = 4 * (int64_t(blockIdx.x) * blockDim.x + threadIdx.x); i < n;
       i += 4 * int64_t(gridDim.x) * blockDim.x) {
    const float4 rv = *reinterpret_cast<const float4*>(r + i);
    const uint32_t sv = solid ? *reinterpret_cast<const uint32_t*>(solid + i) : 0u;
    const float rr[4] = {rv.x, rv.y, rv.z, rv.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if ((sv >> (8 * k)) & 0xffu) continue;
      s += double(rr[k]) * double(rr[k]);
      c += 1.0;
    }
  }
  block_partial(s, ws, q);
  block_partial(c, ws, q + 1);
}

// head: z[v] = fluid ? (sum_c act[v][c] w[c] + b) * out_scale : 0, out_scale = rms(r) * post
__global__ void k_head(int64_t n, const uint16_t* __restrict__ act, const float* __restrict__ w,
                       const float* __restrict__ b, const uint8_t* __restrict__ solid, const double* __restrict__ ws,
                       int q, float post, float* __restrict__ z) {
  float wr[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) wr[c] = w[c];
  const float bb = b[0];
  float out_scale = post;
  if (ws) {
    const double ss = ws[q], cnt = ws[q + 1];
    out_scale = (ss > 0.0 && cnt > 0.0) ? float(sqrt(ss / cnt)) * post : 0.0f;
  }
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const uint4* p = reinterpret_cast<const uint4*>(act + i * 16);
    const uint4 a0 = p[0], a1 = p[1];
    const uint32_t u[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    float s = bb;
#pragma unroll
    for (int c = 0; c < 16; ++c) s += wr[c] * bf16_val(uint16_t((u[c >> 1] >> (16 * (c & 1))) & 0xffffu));
    z[i] = (solid && solid[i]) ? 0.0f : s * out_scale;
  }
}

// head backward: gact[v][c] = gout[v] * w[c]; dw[c] (+)= sum_v gout[v] act[v][c]; db (+)= sum gout
__global__ void k_head_back(int64_t n, const uint16_t* __restrict__ act, const float* __restrict__ gout,
                            const float* __restrict__ w, uint16_t* __restrict__ gact, float* __restrict__ dwb) {
  float wr[16], acc[17];
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    wr[c] = w[c];
    acc[c] = 0.0f;
  }
  acc[16] = 0.0f;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float g = gout[i];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      acc[c] += g * bf16_val(act[i * 16 + c]);
      gact[i * 16 + c] = bf16_bits(g * wr[c]);
    }
    acc[16] += g;
  }
  __shared__ float part[17][4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 17; ++c) {
    float v = acc[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) part[c][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 17) {
    float s = part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] + part[threadIdx.x][3];
    atomicAdd(&dwb[threadIdx.x], s);
  }
}

// masked MSE: gout[v] = fluid ? 2 (out - target) / count : 0 ; ws[q] partial sums of (out - target)^2
// over fluid cells, ws[q+1] the fluid count
__global__ void k_mse(int64_t n, const float* __restrict__ out, const float* __restrict__ target,
                      const uint8_t* __restrict__ solid, float inv_count, float* __restrict__ gout, double* ws,
                      int q) {
  double s = 0.0, c = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const bool fl = !(solid && solid[i]);
    const float d = fl ? out[i] - target[i] : 0.0f;
    gout[i] = 2.0f * d * inv_count;
    s += double(d) * double(d);
    c += fl ? 1.0 : 0.0;
  }
  block_partial(s, ws, q);
  block_partial(c, ws, q + 1);
}

// Adam with bias correction; p, m, v fp32; g fp32
__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// w[27][cin][cout] fp32 -> wt[co][kp] bf16 (forward layout, k = tap*cpack + ci,
// channels cin..cpack-1 zero); flip = 1 packs the data-gradient weights
// wt'[ci][k' = tap'*cout + co] = w[26 - tap'][ci][co] (rows >= cin zero)
__global__ void k_pack_weights(int cin, int cpack, int cout, int rows, int kp, int flip, const float* __restrict__ w,
                               uint16_t* __restrict__ wt) {
  const int64_t n = int64_t(rows) * kp;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const int r = int(i / kp), k = int(i - int64_t(r) * kp);
    float val = 0.0f;
    if (!flip) {
      if (k < 27 * cpack) {
        const int tap = k / cpack, ci = k - tap * cpack;
        if (ci < cin) val = w[(int64_t(tap) * cin + ci) * cout + r];
      }
    } else {
      // r = ci, k = tap'*cout + co
      if (r < cin && k < 27 * cout) {
        const int tp = k / cout, co = k - tp * cout;
        val = w[(int64_t(26 - tp) * cin + r) * cout + co];
      }
    }
    wt[i] = bf16_bits(val);
  }
}

// fp32 -> bf16 copy
__global__ void k_to_bf16(int64_t n, const float* __restrict__ a, uint16_t* __restrict__ b) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    b[i] = bf16_bits(a[i]);
}

static ConvDev conv_dev(const fx_conv_desc* c) {
  ConvDev d;
  d.nx = c->nx; d.ny = c->ny; d.nz = c->nz;
  d.c0 = c->c0; d.c1 = c->c1; d.cin = c->c0 + c->c1; d.cout = c->cout;
  d.kp = (27 * d.cin + 31) / 32 * 32;
  d.up0 = c->up0; d.relu = c->relu;
  return d;
}

static int check_conv(const fx_conv_desc* c) {
  FX_REQUIRE(c, FX_EINVAL, "null conv descriptor");
  FX_REQUIRE(c->nx > 0 && c->ny > 0 && c->nz > 0, FX_EINVAL, "empty conv grid");
  FX_REQUIRE(c->c0 > 0 && c->c0 % 8 == 0 && c->c1 >= 0 && c->c1 % 8 == 0, FX_EINVAL,
             "conv input channels must be multiples of 8");
  FX_REQUIRE(c->cout % 16 == 0 && c->cout >= 16 && c->cout <= 96, FX_EINVAL,
             "conv output channels must be 16..96 in steps of 16");
  FX_REQUIRE(!c->up0 || (c->nx % 2 == 0 && c->ny % 2 == 0 && c->nz % 2 == 0), FX_EINVAL,
             "upsampled input needs an even grid");
  FX_REQUIRE(int64_t(c->nx) * c->ny * c->nz < (int64_t(1) << 31), FX_EINVAL, "conv grid too large");
  return FX_OK;
}

static bool al16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
static bool al4(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 3u) == 0; }

// blocks of 256 threads for `items` work items (grid-stride beyond 16384 blocks)
static dim3 item_grid(int64_t items) {
  return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>((items + 255) / 256, 16384))));
}

int conv_forward(const fx_conv_desc* c, const void* in0, const void* in1, const void* wt, const float* bias,
                 void* out, hipStream_t s) {
  if (int e = check_conv(c)) return e;
  FX_REQUIRE(in0 && wt && out && (c->c1 == 0 || in1), FX_EINVAL, "conv: null array");
  FX_REQUIRE(al16(in0) && al16(wt) && (!in1 || al16(in1)), FX_EINVAL, "conv arrays must be 16-byte aligned");
  ConvDev d = conv_dev(c);
  const int tiles_x = (c->nx + kTX - 1) / kTX, tiles_y = (c->ny + kTY - 1) / kTY, tiles_z = (c->nz + kTZ - 1) / kTZ;
  dim3 grid(tiles_x * tiles_y * tiles_z);
  const uint16_t* a0 = (const uint16_t*)in0;
  const uint16_t* a1 = (const uint16_t*)in1;
  const uint16_t* w = (const uint16_t*)wt;
  uint16_t* o = (uint16_t*)out;
  // the U-Net's forward and data-gradient layers: pipelined kernel, a few
  // z tiles per block (about 1024 blocks or more)
  const int columns = tiles_x * tiles_y;
  const int per_block = std::max(1, std::min(8, tiles_z * columns / 1024));
  const dim3 pgrid(columns * ((tiles_z + per_block - 1) / per_block));
#define FX_PIPE(CI, NT) \
  hipLaunchKernelGGL((k_conv_pipe<CI, NT>), pgrid, dim3(256), 0, s, d, a0, a1, w, bias, o, tiles_x, tiles_y, per_block)
#define FX_CONV(NT) \
  hipLaunchKernelGGL((k_conv_fwd<0, NT>), grid, dim3(256), 0, s, d, a0, a1, w, bias, o, tiles_x, tiles_y)
  const int nt = c->cout / 16, ci = d.cin;
  if (nt == 1 && ci == 8) FX_PIPE(8, 1);
  else if (nt == 1 && ci == 16) FX_PIPE(16, 1);
  else if (nt == 1 && ci == 32) FX_PIPE(32, 1);
  else if (nt == 1 && ci == 48) FX_PIPE(48, 1);
  else if (nt == 2 && ci == 16) FX_PIPE(16, 2);
  else if (nt == 2 && ci == 32) FX_PIPE(32, 2);
  else if (nt == 2 && ci == 64) FX_PIPE(64, 2);
  else if (nt == 2 && ci == 96) FX_PIPE(96, 2);
  else if (nt == 3 && ci == 16) FX_PIPE(16, 3);
  else if (nt == 4 && ci == 32) FX_PIPE(32, 4);
  else if (nt == 4 && ci == 64) FX_PIPE(64, 4);
  else if (nt == 6 && ci == 32) FX_PIPE(32, 6);
  else {
    switch (nt) {
      case 1: FX_CONV(1); break;
      case 2: FX_CONV(2); break;
      case 3: FX_CONV(3); break;
      case 4: FX_CONV(4); break;
      case 5: FX_CONV(5); break;
      default: FX_CONV(6); break;
    }
  }
#undef FX_CONV
#undef FX_PIPE
  FX_LAUNCHED("k_conv_fwd");
  return FX_OK;
}

int unet_forward(const fx_unet* net, const float* r, const uint8_t* solid, double* ws, float post_scale,
                 int output_rms, float* z, hipStream_t s) {
  FX_REQUIRE(net && r && ws && z, FX_EINVAL, "unet forward: null argument");
  FX_REQUIRE(net->nx % 4 == 0 && net->ny % 4 == 0 && net->nz % 4 == 0, FX_EINVAL,
             "unet grid must be divisible by 4");
  for (int i = 0; i < 10; ++i) FX_REQUIRE(net->wt[i] && net->act[i], FX_EINVAL, "unet layer %d missing", i);
  FX_REQUIRE(net->head && net->pool1 && net->pool2 && net->input, FX_EINVAL, "unet buffers missing");
  const int nx = net->nx, ny = net->ny, nz = net->nz;
  const int64_t n = int64_t(nx) * ny * nz;
  auto conv = [&](int layer, int level, int c0, int c1, int cout, int up0, const void* in0, const void* in1) {
    fx_conv_desc d{nx >> level, ny >> level, nz >> level, c0, c1, cout, up0, 1};
    return conv_forward(&d, in0, in1, net->wt[layer], net->bias[layer], net->act[layer], s);
  };
  hipLaunchKernelGGL(k_fluid_norm, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, r, solid, ws, 8);
  FX_LAUNCHED("k_fluid_norm");
  ws_finalize(ws, 8, 2, s);
  hipLaunchKernelGGL(k_unet_input, item_grid(n), dim3(256), 0, s, n, r, solid, (const double*)ws, 8,
                     (uint16_t*)net->input);
  FX_LAUNCHED("k_unet_input");
  if (int e = conv(0, 0, 8, 0, 16, 0, net->input, nullptr)) return e;
  if (int e = conv(1, 0, 16, 0, 16, 0, net->act[0], nullptr)) return e;
  hipLaunchKernelGGL(k_pool, item_grid(n / 8 * 2), dim3(256), 0, s, nx, ny, nz, 16, (const uint16_t*)net->act[1],
                     (uint16_t*)net->pool1);
  if (int e = conv(2, 1, 16, 0, 32, 0, net->pool1, nullptr)) return e;
  if (int e = conv(3, 1, 32, 0, 32, 0, net->act[2], nullptr)) return e;
  hipLaunchKernelGGL(k_pool, item_grid(n / 64 * 4), dim3(256), 0, s, nx >> 1, ny >> 1, nz >> 1, 32,
                     (const uint16_t*)net->act[3], (uint16_t*)net->pool2);
  FX_LAUNCHED("k_pool");
  if (int e = conv(4, 2, 32, 0, 64, 0, net->pool2, nullptr)) return e;
  if (int e = conv(5, 2, 64, 0, 64, 0, net->act[4], nullptr)) return e;
  if (int e = conv(6, 1, 64, 32, 32, 1, net->act[5], net->act[3])) return e;
  if (int e = conv(7, 1, 32, 0, 32, 0, net->act[6], nullptr)) return e;
  if (int e = conv(8, 0, 32, 16, 16, 1, net->act[7], net->act[1])) return e;
  if (int e = conv(9, 0, 16, 0, 16, 0, net->act[8], nullptr)) return e;
  hipLaunchKernelGGL(k_head, item_grid(n), dim3(256), 0, s, n, (const uint16_t*)net->act[9], net->head,
                     net->head + 16, solid, output_rms ? (const double*)ws : nullptr, 8, post_scale, z);
  FX_LAUNCHED("k_head");
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_unet_forward(const fx_unet* net, const float* r, const uint8_t* solid, double* ws, float post_scale,
                    int32_t output_rms, float* z, void* stream) {
  return unet_forward(net, r, solid, ws, post_scale, output_rms, z, S(stream));
}

int fx_conv3d_forward(const fx_conv_desc* conv, const void* in0, const void* in1, const void* wt, const float* bias,
                      void* out, void* stream) {
  return conv_forward(conv, in0, in1, wt, bias, out, S(stream));
}

int fx_conv3d_weight_grad(const fx_conv_desc* conv, const void* in0, const void* in1, const void* g,
                          const void* act, void* work, int64_t work_bytes, float* dw, float* db, int32_t accumulate,
                          void* stream) {
  if (int e = check_conv(conv)) return e;
  FX_REQUIRE(in0 && g && dw && work && (conv->c1 == 0 || in1), FX_EINVAL, "weight grad: null array");
  FX_REQUIRE(conv->nx % 8 == 0, FX_EINVAL, "weight grad needs nx % 8 == 0");
  FX_REQUIRE(conv->cout <= 64, FX_EINVAL, "weight grad supports up to 64 output channels");
  ConvDev d = conv_dev(conv);
  hipStream_t s = S(stream);
  const int nvox = conv->nx * conv->ny * conv->nz;
  const int64_t ngroups = nvox / 8;
  // chunks: enough waves to fill the chip, partial buffer within the workspace
  const int ci_tiles = (d.cin + 15) / 16;
  int chunks = int(std::max<int64_t>(1, std::min<int64_t>(ngroups / 64, 4096 / (9 * ci_tiles) + 1)));
  const int64_t wn = int64_t(27) * d.cin * d.cout;
  const int64_t t_bytes = (int64_t(d.cin) + d.cout) * nvox * 2;
  const int64_t t_bytes_al = (t_bytes + 255) / 256 * 256;
  while (chunks > 1 && t_bytes_al + int64_t(chunks) * wn * 4 > work_bytes) chunks /= 2;
  FX_REQUIRE(t_bytes_al + int64_t(chunks) * wn * 4 <= work_bytes, FX_EINVAL,
             "weight grad workspace too small: need %lld bytes", (long long)(t_bytes_al + wn * 4));
  int64_t gpc = (ngroups + chunks - 1) / chunks;
  gpc = (gpc + 3) / 4 * 4;
  uint16_t* inT = (uint16_t*)work;
  uint16_t* gT = inT + int64_t(d.cin) * nvox;
  float* part = (float*)((char*)work + t_bytes_al);
  hipLaunchKernelGGL(k_transpose_in, dim3((nvox + 255) / 256), dim3(256), 0, s, d, (const uint16_t*)in0,
                     (const uint16_t*)in1, inT, nvox);
  hipLaunchKernelGGL(k_transpose_grad, dim3((nvox + 255) / 256), dim3(256), 0, s, nvox, d.cout, (const uint16_t*)g,
                     (const uint16_t*)act, gT);
  FX_LAUNCHED("weight grad transposes");
  const int ntask = 9 * ci_tiles * chunks;
  dim3 grid((ntask + 3) / 4);
  switch (d.cout / 16) {
    case 1: hipLaunchKernelGGL(k_conv_wgrad<1>, grid, dim3(256), 0, s, d.nx, d.ny, d.nz, d.cin, d.cout, inT, gT, part, chunks, int(gpc), ci_tiles); break;
    case 2: hipLaunchKernelGGL(k_conv_wgrad<2>, grid, dim3(256), 0, s, d.nx, d.ny, d.nz, d.cin, d.cout, inT, gT, part, chunks, int(gpc), ci_tiles); break;
    case 3: hipLaunchKernelGGL(k_conv_wgrad<3>, grid, dim3(256), 0, s, d.nx, d.ny, d.nz, d.cin, d.cout, inT, gT, part, chunks, int(gpc), ci_tiles); break;
    default: hipLaunchKernelGGL(k_conv_wgrad<4>, grid, dim3(256), 0, s, d.nx, d.ny, d.nz, d.cin, d.cout, inT, gT, part, chunks, int(gpc), ci_tiles); break;
  }
  FX_LAUNCHED("k_conv_wgrad");
  hipLaunchKernelGGL(k_sum_chunks, dim3(int((wn + 255) / 256)), dim3(256), 0, s, wn, chunks, part, dw, accumulate);
  FX_LAUNCHED("k_sum_chunks");
  if (db) {
    hipLaunchKernelGGL(k_bias_grad, dim3(d.cout), dim3(256), 0, s, nvox, d.cout, (const uint16_t*)g,
                       (const uint16_t*)act, db, accumulate);
    FX_LAUNCHED("k_bias_grad");
  }
  return FX_OK;
}

int64_t fx_conv3d_weight_grad_workspace(const fx_conv_desc* conv) {
  if (check_conv(conv)) return -1;
  ConvDev d = conv_dev(conv);
  const int64_t nvox = int64_t(conv->nx) * conv->ny * conv->nz;
  const int64_t t_bytes = ((int64_t(d.cin) + d.cout) * nvox * 2 + 255) / 256 * 256;
  const int ci_tiles = (d.cin + 15) / 16;
  const int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(nvox / 8 / 64, 4096 / (9 * ci_tiles) + 1));
  return t_bytes + chunks * 27 * d.cin * d.cout * 4;
}

int fx_conv_pack_weights(int32_t cin, int32_t cpack, int32_t cout, int32_t flip, const float* w, void* wt,
                         void* stream) {
  FX_REQUIRE(cin > 0 && cpack >= cin && cout > 0 && w && wt, FX_EINVAL, "pack weights: bad arguments");
  int rows, kp;
  if (!flip) {
    rows = cout;
    kp = (27 * cpack + 31) / 32 * 32;
  } else {
    rows = (cin + 15) / 16 * 16;
    kp = (27 * cout + 31) / 32 * 32;
  }
  const int64_t n = int64_t(rows) * kp;
  hipLaunchKernelGGL(k_pack_weights, dim3(int((n + 255) / 256)), dim3(256), 0, S(stream), cin, cpack, cout, rows, kp,
                     flip, w, (uint16_t*)wt);
  FX_LAUNCHED("k_pack_weights");
  return FX_OK;
}

int fx_relu_mask(int64_t n, void* g, const void* act, void* stream) {
  FX_REQUIRE(n >= 0 && g && act, FX_EINVAL, "relu mask: bad arguments");
  hipLaunchKernelGGL(k_relu_mask, dim3(2048), dim3(256), 0, S(stream), n, (uint16_t*)g, (const uint16_t*)act);
  FX_LAUNCHED("k_relu_mask");
  return FX_OK;
}

int fx_pool(int32_t nx, int32_t ny, int32_t nz, int32_t c, const void* in, void* out, void* stream) {
  FX_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0 && c > 0 && c % 8 == 0 && in && out && al16(in) && al16(out),
             FX_EINVAL, "pool: bad arguments (even grid, channels a multiple of 8, 16-byte aligned arrays)");
  hipLaunchKernelGGL(k_pool, item_grid(int64_t(nx / 2) * (ny / 2) * (nz / 2) * (c / 8)), dim3(256), 0, S(stream), nx,
                     ny, nz, c, (const uint16_t*)in,
                     (uint16_t*)out);
  FX_LAUNCHED("k_pool");
  return FX_OK;
}

int fx_pool_backward_add(int32_t nx, int32_t ny, int32_t nz, int32_t c, const void* gpool, const void* skip,
                         void* out, void* stream) {
  FX_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0 && c > 0 && gpool && out, FX_EINVAL,
             "pool backward: bad arguments");
  hipLaunchKernelGGL(k_pool_back_add, dim3(2048), dim3(256), 0, S(stream), nx, ny, nz, c, (const uint16_t*)gpool,
                     (const uint16_t*)skip, (uint16_t*)out);
  FX_LAUNCHED("k_pool_back_add");
  return FX_OK;
}

int fx_concat_backward(int32_t nx, int32_t ny, int32_t nz, int32_t c0, int32_t c1, const void* g, void* gup,
                       void* gskip, void* stream) {
  FX_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0 && c0 > 0 && c1 > 0 && g && gup && gskip, FX_EINVAL,
             "concat backward: bad arguments");
  hipLaunchKernelGGL(k_concat_back, dim3(2048), dim3(256), 0, S(stream), nx, ny, nz, c0, c1, (const uint16_t*)g,
                     (uint16_t*)gup, (uint16_t*)gskip);
  FX_LAUNCHED("k_concat_back");
  return FX_OK;
}

int fx_unet_input(int64_t n, const float* r, const uint8_t* solid, double* ws, void* out, void* stream) {
  FX_REQUIRE(n > 0 && n % 4 == 0 && r && ws && out && al16(out) && al16(r) && (!solid || al4(solid)), FX_EINVAL,
             "unet input: bad arguments (n a multiple of 4, 16-byte aligned r and out)");
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_fluid_norm, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, r, solid, ws, 8);
  FX_LAUNCHED("k_fluid_norm");
  ws_finalize(ws, 8, 2, s);
  hipLaunchKernelGGL(k_unet_input, item_grid(n), dim3(256), 0, s, n, r, solid, (const double*)ws, 8,
                     (uint16_t*)out);
  FX_LAUNCHED("k_unet_input");
  return FX_OK;
}

int fx_unet_head(int64_t n, const void* act, const float* w, const float* b, const uint8_t* solid, const double* ws,
                 float post_scale, float* z, void* stream) {
  FX_REQUIRE(n > 0 && n % 4 == 0 && act && w && b && z && al16(act) && al16(z) && (!solid || al4(solid)), FX_EINVAL,
             "unet head: bad arguments (n a multiple of 4, 16-byte aligned act and z)");
  hipLaunchKernelGGL(k_head, item_grid(n), dim3(256), 0, S(stream), n, (const uint16_t*)act, w, b, solid, ws, 8,
                     post_scale, z);
  FX_LAUNCHED("k_head");
  return FX_OK;
}

int fx_unet_head_backward(int64_t n, const void* act, const float* gout, const float* w, void* gact, float* dwb,
                          void* stream) {
  FX_REQUIRE(n > 0 && act && gout && w && gact && dwb, FX_EINVAL, "unet head backward: bad arguments");
  hipLaunchKernelGGL(k_head_back, dim3(1024), dim3(256), 0, S(stream), n, (const uint16_t*)act, gout, w,
                     (uint16_t*)gact, dwb);
  FX_LAUNCHED("k_head_back");
  return FX_OK;
}

int fx_masked_mse(int64_t n, const float* out, const float* target, const uint8_t* solid, float inv_count,
                  float* gout, double* ws, void* stream) {
  FX_REQUIRE(n > 0 && out && target && gout && ws, FX_EINVAL, "mse: bad arguments");
  hipStream_t s = S(stream);
  hipLaunchKernelGGL(k_mse, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, out, target, solid, inv_count, gout, ws,
                     10);
  FX_LAUNCHED("k_mse");
  ws_finalize(ws, 10, 2, s);
  return FX_OK;
}

int fx_adam(int64_t n, float* p, const float* g, float* m, float* v, int32_t t, float lr, float beta1, float beta2,
            float eps, void* stream) {
  FX_REQUIRE(n > 0 && p && g && m && v && t >= 1, FX_EINVAL, "adam: bad arguments");
  const float c1 = 1.0f - powf(beta1, float(t)), c2 = 1.0f - powf(beta2, float(t));
  hipLaunchKernelGGL(k_adam, dim3(int(std::min<int64_t>((n + 255) / 256, 4096))), dim3(256), 0, S(stream), n, p, g,
                     m, v, lr, beta1, beta2, eps, c1, c2);
  FX_LAUNCHED("k_adam");
  return FX_OK;
}

int fx_to_bf16(int64_t n, const float* a, void* b, void* stream) {
  FX_REQUIRE(n >= 0 && a && b, FX_EINVAL, "to_bf16: bad arguments");
  hipLaunchKernelGGL(k_to_bf16, dim3(2048), dim3(256), 0, S(stream), n, a, (uint16_t*)b);
  FX_LAUNCHED("k_to_bf16");
  return FX_OK;
}

}  // extern "C"
