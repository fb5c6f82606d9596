// Kernel (a): macrocell value-range grid and majorant grid.
//
// Reference: renderer.cu:1116-1179 buildSpatialMacrocells scatters, per voxel
// (one thread each, one launch per channel), the min/max of its 3x3x3
// neighbourhood into every macrocell its widened box projects to, with CAS
// float atomics. Here the same grid is gathered instead: the host derives, per
// axis, the window of voxels whose projected boxes cover each macrocell index
// (an interval, because both projected box ends grow monotonically with the
// voxel index) and widens it by the neighbourhood; one workgroup per
// (macrocell, channel) reduces its window with wave64 shuffles. No atomics,
// one launch for all channels, and every voxel byte is read from HBM in
// coalesced rows along x.
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
  const int total = lx * ly * lz;
  const uint8_t* base = P.vol + size_t(ch) * P.chan_stride;
  const uint32_t nx = uint32_t(P.nx), nxy = uint32_t(P.nx) * uint32_t(P.ny);
  float vmin = __int_as_float(0x7f800000), vmax = -__int_as_float(0x7f800000);
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int ix = i % lx;
    const int r = i / lx;
    const int iy = r % ly;
    const int iz = r / ly;
    const uint32_t idx = uint32_t(x0 + ix) + uint32_t(y0 + iy) * nx + uint32_t(z0 + iz) * nxy;
    const float v = fetch_raw<DT>(base, idx);
    vmin = fminf(vmin, v);
    vmax = fmaxf(vmax, v);
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
      m = channel_majorant(tab + size_t(c) * P.tf.stride, P.tf.n[c], P.tf.opacity[c], P.tf.xf_lo[c],
                           P.tf.xf_hi[c], P.tf.dom_lo[c], P.tf.dom_hi[c], r.x, r.y);
    P.maj[i] = m;
  } else {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.n_cells) return;
    float sum = 0.f;
    for (int c = 0; c < C; ++c) {
      const float2 r = rg[size_t(cell) * C + c];
      if (r.y < r.x) continue;
      sum += channel_majorant(tab + size_t(c) * P.tf.stride, P.tf.n[c], P.tf.opacity[c],
                              P.tf.xf_lo[c], P.tf.xf_hi[c], P.tf.dom_lo[c], P.tf.dom_hi[c], r.x,
                              r.y);
    }
    P.maj[cell] = sum;
  }
}

}  // namespace

hipError_t launch_ranges(const RangeParams& p, int dtype, hipStream_t s) {
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

hipError_t launch_majorants(const MajorantParams& p, hipStream_t s) {
  const int work = p.mode <= 2 ? p.n_cells * p.channels : p.n_cells;
  const int blocks = (work + 255) / 256;
  hipLaunchKernelGGL(majorant_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace mdwt
