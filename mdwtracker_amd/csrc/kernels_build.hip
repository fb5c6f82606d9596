// Kernel (a): macrocell value-range grid and majorant grid.
//
// Reference: renderer.cu:1116-1179 buildSpatialMacrocells scatters, per voxel
// (one thread each, one launch per channel), the min/max of its 3x3x3
// neighbourhood into every macrocell its widened box projects to, with CAS
// float atomics. Here the same grid is gathered instead: the host derives, per
// axis, the window of voxels whose projected boxes cover each macrocell index
// (an interval, because both projected box ends grow monotonically with the
// voxel index) and widens it by the neighbourhood. Because the window is a
// box, its min/max separates: range_pencil_kernel gives one workgroup to a
// pencil of cells along x (fixed y and z windows), reads those rows in full
// with 16-byte loads, keeps per-column min/max and reduces each cell's x
// window from LDS. Volumes whose rows are not a multiple of 16 bytes use
// range_kernel, one workgroup per (macrocell, channel) over its window in
// dword loads. No atomics, one launch for all channels.
#include "mdwt_device.hpp"
#include "mdwt_internal.hpp"

namespace mdwt {
namespace {

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

template <int DT>
__host__ __device__ constexpr int voxel_bytes() {
  return DT == MDWT_DT_U8 ? 1 : DT == MDWT_DT_F32 ? 4 : 2;
}
// Voxel e of a little-endian dword holding 4 / voxel_bytes voxels, as
// fetch_raw would convert it.
template <int DT>
__device__ __forceinline__ float word_voxel(uint32_t w, int e) {
  if constexpr (DT == MDWT_DT_U8) return float((w >> (8 * e)) & 0xffu);
  if constexpr (DT == MDWT_DT_U16) return float((w >> (16 * e)) & 0xffffu);
  if constexpr (DT == MDWT_DT_F16) {
    const uint16_t h = uint16_t(w >> (16 * e));
    return float(__builtin_bit_cast(_Float16, h));
  }
  return __uint_as_float(w);
}

template <int DT>
__global__ __launch_bounds__(256) void range_kernel(RangeParams P) {
  const int cell = blockIdx.x;
  const int ch = blockIdx.y;
  const int mx = P.mc[0], my = P.mc[1];
  const int cx = cell % mx;
  const int cy = (cell / mx) % my;
  const int cz = cell / (mx * my);
  const int* wx = P.win;
  const int* wy = P.win + 2 * P.mc[0];
  const int* wz = P.win + 2 * (P.mc[0] + P.mc[1]);
  const int x0 = wx[2 * cx], x1 = wx[2 * cx + 1];
  const int y0 = wy[2 * cy], y1 = wy[2 * cy + 1];
  const int z0 = wz[2 * cz], z1 = wz[2 * cz + 1];
  float2* out = reinterpret_cast<float2*>(P.ranges);
  if (x0 > x1 || y0 > y1 || z0 > z1) {  // no voxel covers this cell
    if (threadIdx.x == 0) out[size_t(cell) * P.channels + ch] = make_float2(1e20f, -1e20f);
    return;
  }
  const int lx = x1 - x0 + 1, ly = y1 - y0 + 1, lz = z1 - z0 + 1;
  const uint8_t* base = P.vol + size_t(ch) * P.chan_stride;
  float vmin = __int_as_float(0x7f800000), vmax = -__int_as_float(0x7f800000);
  // One row segment (lx voxels along x at fixed y, z) spans at most ndw
  // aligned dwords; gl lanes (a power of two >= ndw, at most 64) share a row
  // and rp = 256 / gl rows are read per pass, four passes' loads issued before
  // any is reduced. Rows advance by additions (iy, iz carried), no per-voxel
  // index division. A dword may hold voxels outside the window: they are
  // masked by their byte position. The channel's allocation is padded to 256
  // bytes, so the dword holding its last voxel is inside it.
  constexpr int ES = voxel_bytes<DT>();
  const int span = lx * ES;
  const int ndw = (span + 6) >> 2;
  int lg = 0;
  while ((1 << lg) < ndw && lg < 6) ++lg;
  const int gl = 1 << lg, rp = int(blockDim.x) >> lg;
  const int q = int(threadIdx.x) & (gl - 1), g = int(threadIdx.x) >> lg;
  const int nrows = ly * lz;
  const int dy = rp % ly, dz = rp / ly;
  const uint32_t nx = uint32_t(P.nx), ny = uint32_t(P.ny);
  for (int k0 = 0; k0 < ndw; k0 += gl) {
    const int k = k0 + q;
    int r = g, iy = g % ly, iz = g / ly;
    while (r < nrows) {
      uint32_t w[4];
      int s[4];
      bool ok[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ok[u] = false;
        s[u] = 0;
        w[u] = 0u;
        if (r < nrows) {
          const size_t row = (size_t(uint32_t(z0 + iz)) * ny + uint32_t(y0 + iy)) * nx + uint32_t(x0);
          const uintptr_t lo = reinterpret_cast<uintptr_t>(base) + row * ES;
          s[u] = int(lo & 3u);
          ok[u] = k < ndw && 4 * k <= s[u] + span - 1;
          if (ok[u]) w[u] = *reinterpret_cast<const uint32_t*>((lo & ~uintptr_t(3)) + 4u * uint32_t(k));
          r += rp;
          iy += dy;
          iz += dz;
          if (iy >= ly) {
            iy -= ly;
            ++iz;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (!ok[u]) continue;
        const int end = s[u] + span - 1;
#pragma unroll
        for (int e = 0; e < 4 / ES; ++e) {
          const int pos = 4 * k + e * ES;
          if (pos < s[u] || pos > end) continue;
          const float v = word_voxel<DT>(w[u], e);
          vmin = fminf(vmin, v);
          vmax = fmaxf(vmax, v);
        }
      }
    }
  }
  vmin = wave_min(vmin);
  vmax = wave_max(vmax);
  __shared__ float smin[4], smax[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    smin[wave] = vmin;
    smax[wave] = vmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = smin[0], b = smax[0];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) {
      a = fminf(a, smin[w]);
      b = fmaxf(b, smax[w]);
    }
    // value conversion is monotone and exact on the integer types, so the
    // converted min/max equal min/max of converted voxels
    out[size_t(cell) * P.channels + ch] = make_float2(a * P.scale + P.offset, b * P.scale + P.offset);
  }
}

// Same grid, gathered per pencil: one workgroup per (cy, cz, channel) reads
// every voxel row of the pencil's y and z windows in full, as aligned 16-byte
// chunks (J chunks per lane, gl lanes per row), keeps a running min/max per x
// column in registers, merges the row slots by shuffles and the four waves in
// LDS, then reduces each cell's x window from the LDS columns. A voxel row is
// read once per (cy, cz) instead of once per cell, and each lane issues U
// chunk loads before reducing any. Requires rows of a multiple of 16 bytes
// (and at most kPencilChunks chunks); other volumes use range_kernel.
constexpr int kPencilChunks = 128;  // 16-byte chunks per row
constexpr int kPencilColumns = 2048;

template <int DT>
__device__ __forceinline__ float chunk_voxel(const uint4& c, int v) {
  const uint32_t w[4] = {c.x, c.y, c.z, c.w};
  if constexpr (DT == MDWT_DT_U8) return word_voxel<DT>(w[v >> 2], v & 3);
  if constexpr (DT == MDWT_DT_U16 || DT == MDWT_DT_F16) return word_voxel<DT>(w[v >> 1], v & 1);
  return word_voxel<DT>(w[v], 0);
}

template <int DT, int J>
__global__ __launch_bounds__(256) void range_pencil_kernel(RangeParams P) {
  constexpr int ES = voxel_bytes<DT>(), VPC = 16 / ES, U = J == 1 ? 8 : 4;
  __shared__ float cmin[kPencilColumns], cmax[kPencilColumns];
  const int cy = blockIdx.x, cz = blockIdx.y, ch = blockIdx.z;
  const int mx = P.mc[0], my = P.mc[1];
  const int* wx = P.win;
  const int* wy = P.win + 2 * P.mc[0];
  const int* wz = P.win + 2 * (P.mc[0] + P.mc[1]);
  const int y0 = wy[2 * cy], y1 = wy[2 * cy + 1];
  const int z0 = wz[2 * cz], z1 = wz[2 * cz + 1];
  float2* out = reinterpret_cast<float2*>(P.ranges) + (size_t(cz) * my + cy) * mx * P.channels + ch;
  if (y0 > y1 || z0 > z1) {  // no voxel covers any cell of the pencil
    for (int cx = threadIdx.x; cx < mx; cx += blockDim.x)
      out[size_t(cx) * P.channels] = make_float2(1e20f, -1e20f);
    return;
  }
  const int ly = y1 - y0 + 1, lz = z1 - z0 + 1, nrows = ly * lz;
  const int chunks = P.nx * ES / 16;
  int lg = 0;
  while ((1 << lg) * J < chunks && lg < 6) ++lg;
  const int gl = 1 << lg, rp = int(blockDim.x) >> lg;
  const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
  const int q = int(threadIdx.x) & (gl - 1), g = int(threadIdx.x) >> lg;
  const uint8_t* base = P.vol + size_t(ch) * P.chan_stride;
  const size_t row_bytes = size_t(P.nx) * ES;
  float vmin[J][VPC], vmax[J][VPC];
#pragma unroll
  for (int j = 0; j < J; ++j)
#pragma unroll
    for (int v = 0; v < VPC; ++v) {
      vmin[j][v] = __int_as_float(0x7f800000);
      vmax[j][v] = -__int_as_float(0x7f800000);
    }
  const int dy = rp % ly, dz = rp / ly;
  int r = g, iy = g % ly, iz = g / ly;
  while (r < nrows) {
    uint4 w[U][J];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ok[u] = r < nrows;
      if (ok[u]) {
        const uint8_t* row = base + (size_t(z0 + iz) * P.ny + size_t(y0 + iy)) * row_bytes;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const int c = q + j * gl;
          w[u][j] = c < chunks ? *reinterpret_cast<const uint4*>(row + size_t(c) * 16) : uint4{};
        }
        r += rp;
        iy += dy;
        iz += dz;
        if (iy >= ly) {
          iy -= ly;
          ++iz;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!ok[u]) continue;
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int v = 0; v < VPC; ++v) {
          const float x = chunk_voxel<DT>(w[u][j], v);
          vmin[j][v] = fminf(vmin[j][v], x);
          vmax[j][v] = fmaxf(vmax[j][v], x);
        }
    }
  }
  // lanes q, q + gl, ... of a wave hold the same columns (other rows)
  for (int o = gl; o < 64; o <<= 1)
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int v = 0; v < VPC; ++v) {
        vmin[j][v] = fminf(vmin[j][v], __shfl_xor(vmin[j][v], o, 64));
        vmax[j][v] = fmaxf(vmax[j][v], __shfl_xor(vmax[j][v], o, 64));
      }
  const int waves = int(blockDim.x) >> 6;
  for (int wv = 0; wv < waves; ++wv) {
    if (wave == wv && lane < gl) {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int c = q + j * gl;
        if (c >= chunks) continue;
#pragma unroll
        for (int v = 0; v < VPC; ++v) {
          const int x = c * VPC + v;
          cmin[x] = wv == 0 ? vmin[j][v] : fminf(cmin[x], vmin[j][v]);
          cmax[x] = wv == 0 ? vmax[j][v] : fmaxf(cmax[x], vmax[j][v]);
        }
      }
    }
    __syncthreads();
  }
  // 16 lanes per cell reduce its x window from the columns
  const int sub = int(threadIdx.x) & 15;
  for (int cx0 = 0; cx0 < mx; cx0 += int(blockDim.x) >> 4) {
    const int cx = cx0 + (int(threadIdx.x) >> 4);
    const bool live = cx < mx;
    const int x0 = live ? wx[2 * cx] : 0, x1 = live ? wx[2 * cx + 1] : -1;
    float a = __int_as_float(0x7f800000), b = -__int_as_float(0x7f800000);
    for (int x = x0 + sub; x <= x1; x += 16) {
      a = fminf(a, cmin[x]);
      b = fmaxf(b, cmax[x]);
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) {
      a = fminf(a, __shfl_xor(a, o, 64));
      b = fmaxf(b, __shfl_xor(b, o, 64));
    }
    if (live && sub == 0)
      out[size_t(cx) * P.channels] = x0 > x1 ? make_float2(1e20f, -1e20f)
                                             : make_float2(a * P.scale + P.offset,
                                                           b * P.scale + P.offset);
  }
}

// Per-channel majorant of one cell, renderer.cu:159-203 (the texel-interval
// max plus the two extra lerp lookups at addr + 0.5).
__device__ __forceinline__ float channel_majorant(const float4* tab, int n, float os, float xlo,
                                                  float xhi, float vlo, float vhi, float mn,
                                                  float mx) {
  const float remappedMin1 = (mn - vlo) / (vhi - vlo);
  const float remappedMin = (remappedMin1 - xlo) / (xhi - xlo);
  const float remappedMax1 = (mx - vlo) / (vhi - vlo);
  const float remappedMax = (remappedMax1 - xlo) / (xhi - xlo);
  const float addr1 = remappedMin * float(n);
  const float addr2 = remappedMax * float(n);
  const int addrMin = clamp_index(fminf(floorf(addr1), floorf(addr2)), n);
  const int addrMax = clamp_index(fmaxf(ceilf(addr1), ceilf(addr2)), n);
  float m = 0.f;
  for (int i = addrMin; i <= addrMax; ++i) {
    const float d = tf_lerp_alpha(tab, n, (float(i) + 0.5f) / float(n)) * os;
    m = (i == addrMin) ? d : fmaxf(m, d);
  }
  float d = tf_lerp_alpha(tab, n, (addr1 + 0.5f) / float(n)) * os;
  m = fmaxf(m, d);
  d = tf_lerp_alpha(tab, n, (addr2 + 0.5f) / float(n)) * os;
  m = fmaxf(m, d);
  return m;
}

// An option beside renderer.cu:159-203, whose majorant can fall below the
// opacities a lookup interpolates by up to half a texel of lerp: the largest
// opacity of every texel that a lookup of a value in [mn, mx] reads. The
// lookup interval is widened by 2^-12 (1 + |r|) on each side, which covers the
// render's remap by reciprocals against the divisions here, and the maximum
// is scaled by 1 + 2^-20 for the lerp's rounding. A lookup at u reads
// texels floor(u n - 0.5) and the next one, clamped to the table.
__device__ __forceinline__ float channel_majorant_conservative(const float4* tab, int n, float os,
                                                               float xlo, float xhi, float vlo,
                                                               float vhi, float mn, float mx) {
  const float r1 = ((mn - vlo) / (vhi - vlo) - xlo) / (xhi - xlo);
  const float r2 = ((mx - vlo) / (vhi - vlo) - xlo) / (xhi - xlo);
  const float a = fminf(r1, r2), b = fmaxf(r1, r2);
  const float lo = a - 0x1p-12f * (1.0f + fabsf(a)), hi = b + 0x1p-12f * (1.0f + fabsf(b));
  const int i0 = clamp_index(floorf(lo * float(n) - 0.5f), n);
  const int i1 = clamp_index_plus1(floorf(hi * float(n) - 0.5f), n);
  float m = 0.f;
  for (int i = i0; i <= i1; ++i) m = fmaxf(m, tab[i].w);
  return m * os * (1.0f + 0x1p-20f);
}

// renderer.cu:206-274: modes 0/2 keep one majorant per (cell, channel)
// (renderer.cu:159-203), mode 3 their sum per cell (renderer.cu:111-157).
__global__ __launch_bounds__(256) void majorant_kernel(MajorantParams P) {
  const int C = P.channels;
  const float4* tab = reinterpret_cast<const float4*>(P.tf.tab);
  const float2* rg = reinterpret_cast<const float2*>(P.ranges);
  if (P.mode <= 2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n_cells * C) return;
    const int c = i % C;
    const float2 r = rg[i];
    float m = 0.f;
    if (!(r.y < r.x))
      m = P.conservative
              ? channel_majorant_conservative(tab + size_t(c) * P.tf.stride, P.tf.n[c],
                                              P.tf.opacity[c], P.tf.xf_lo[c], P.tf.xf_hi[c],
                                              P.tf.dom_lo[c], P.tf.dom_hi[c], r.x, r.y)
              : channel_majorant(tab + size_t(c) * P.tf.stride, P.tf.n[c], P.tf.opacity[c],
                                 P.tf.xf_lo[c], P.tf.xf_hi[c], P.tf.dom_lo[c], P.tf.dom_hi[c], r.x,
                                 r.y);
    P.maj[i] = m;
  } else {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.n_cells) return;
    float sum = 0.f;
    for (int c = 0; c < C; ++c) {
      const float2 r = rg[size_t(cell) * C + c];
      if (r.y < r.x) continue;
      sum += P.conservative
                 ? channel_majorant_conservative(tab + size_t(c) * P.tf.stride, P.tf.n[c],
                                                 P.tf.opacity[c], P.tf.xf_lo[c], P.tf.xf_hi[c],
                                                 P.tf.dom_lo[c], P.tf.dom_hi[c], r.x, r.y)
                 : channel_majorant(tab + size_t(c) * P.tf.stride, P.tf.n[c], P.tf.opacity[c],
                                    P.tf.xf_lo[c], P.tf.xf_hi[c], P.tf.dom_lo[c], P.tf.dom_hi[c],
                                    r.x, r.y);
    }
    P.maj[cell] = sum;
  }
}

// Brick copy of the volume (the samplers' layout, VolumeParams): output
// voxel o of a channel is voxel (o & 3, (o >> 2) & 3, o >> 4) of brick
// o >> 6 (bricks x fastest); voxels past the volume's edge are zero (the
// samplers clamp their indices, so they are never read). One thread per
// output voxel, grid (bricks * 64 / 256, channels).
template <typename T>
__global__ __launch_bounds__(256) void brick_kernel(const uint8_t* __restrict__ vol, uint64_t chan_stride,
                                                    int nx, int ny, int nz, int nbx, int nby,
                                                    uint64_t per, uint8_t* __restrict__ bricks,
                                                    uint64_t brick_stride) {
  const uint64_t o = uint64_t(blockIdx.x) * 256u + threadIdx.x;
  if (o >= per) return;
  const uint32_t b = uint32_t(o >> 6), w = uint32_t(o & 63u);
  const uint32_t bx = b % uint32_t(nbx), r = b / uint32_t(nbx);
  const uint32_t by = r % uint32_t(nby), bz = r / uint32_t(nby);
  const uint32_t x = bx * 4u + (w & 3u), y = by * 4u + ((w >> 2) & 3u), z = bz * 4u + (w >> 4);
  const T* src = reinterpret_cast<const T*>(vol + size_t(blockIdx.y) * chan_stride);
  T* dst = reinterpret_cast<T*>(bricks + size_t(blockIdx.y) * brick_stride);
  dst[o] = (x < uint32_t(nx) && y < uint32_t(ny) && z < uint32_t(nz))
               ? src[(size_t(z) * uint32_t(ny) + y) * uint32_t(nx) + x]
               : T(0);
}

template <int DT>
void launch_pencils(const RangeParams& p, hipStream_t s) {
  const dim3 grid(unsigned(p.mc[1]), unsigned(p.mc[2]), unsigned(p.channels));
  if (p.nx * voxel_bytes<DT>() <= 64 * 16)
    hipLaunchKernelGGL((range_pencil_kernel<DT, 1>), grid, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((range_pencil_kernel<DT, 2>), grid, dim3(256), 0, s, p);
}

int dtype_bytes(int dtype) { return dtype == MDWT_DT_U8 ? 1 : dtype == MDWT_DT_F32 ? 4 : 2; }

}  // namespace

hipError_t launch_ranges(const RangeParams& p, int dtype, hipStream_t s) {
  const int row = p.nx * dtype_bytes(dtype);
  if (row % 16 == 0 && row / 16 <= kPencilChunks) {
    switch (dtype) {
      case MDWT_DT_U8: launch_pencils<MDWT_DT_U8>(p, s); break;
      case MDWT_DT_U16: launch_pencils<MDWT_DT_U16>(p, s); break;
      case MDWT_DT_F16: launch_pencils<MDWT_DT_F16>(p, s); break;
      default: launch_pencils<MDWT_DT_F32>(p, s); break;
    }
    return hipGetLastError();
  }
  const int n_cells = p.mc[0] * p.mc[1] * p.mc[2];
  const dim3 grid(unsigned(n_cells), unsigned(p.channels));
  switch (dtype) {
    case MDWT_DT_U8: hipLaunchKernelGGL(range_kernel<MDWT_DT_U8>, grid, dim3(256), 0, s, p); break;
    case MDWT_DT_U16: hipLaunchKernelGGL(range_kernel<MDWT_DT_U16>, grid, dim3(256), 0, s, p); break;
    case MDWT_DT_F16: hipLaunchKernelGGL(range_kernel<MDWT_DT_F16>, grid, dim3(256), 0, s, p); break;
    default: hipLaunchKernelGGL(range_kernel<MDWT_DT_F32>, grid, dim3(256), 0, s, p); break;
  }
  return hipGetLastError();
}

hipError_t launch_bricks(const VolumeParams& v, int dtype, hipStream_t s) {
  const uint64_t per = uint64_t(v.nbx) * uint64_t(v.nby) * uint64_t((v.nz + 3) / 4) * 64u;
  const dim3 grid(unsigned((per + 255) / 256), unsigned(v.channels));
  switch (dtype_bytes(dtype)) {
    case 1:
      hipLaunchKernelGGL(brick_kernel<uint8_t>, grid, dim3(256), 0, s, v.vol, v.chan_stride, v.nx, v.ny,
                         v.nz, v.nbx, v.nby, per, const_cast<uint8_t*>(v.bricks), v.brick_stride);
      break;
    case 2:
      hipLaunchKernelGGL(brick_kernel<uint16_t>, grid, dim3(256), 0, s, v.vol, v.chan_stride, v.nx, v.ny,
                         v.nz, v.nbx, v.nby, per, const_cast<uint8_t*>(v.bricks), v.brick_stride);
      break;
    default:
      hipLaunchKernelGGL(brick_kernel<uint32_t>, grid, dim3(256), 0, s, v.vol, v.chan_stride, v.nx, v.ny,
                         v.nz, v.nbx, v.nby, per, const_cast<uint8_t*>(v.bricks), v.brick_stride);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_majorants(const MajorantParams& p, hipStream_t s) {
  const int work = p.mode <= 2 ? p.n_cells * p.channels : p.n_cells;
  const int blocks = (work + 255) / 256;
  hipLaunchKernelGGL(majorant_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace mdwt
