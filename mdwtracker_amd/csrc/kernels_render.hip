// Kernels (b)+(c): per-pixel ray generation, root clip, 3D-DDA over the
// majorant grid and the per-channel Woodcock tracking loop, for -m 0 (N-DDA),
// -m 2 (syncDDA) and -m 3 (cumulative majorant), plus accumulation.
//
// Reference: mainRG (deviceCode.cu:283-479) launched as one OptiX thread per
// pixel, volumeIntersection (deviceCode.cu:1676-1726), the closest-hit
// programs baseLineDTCH (858-970), altMultiMajDTCH (747-856) and
// cummilativeDTCH (513-617) running dda::dda3 (dda.h:82-203) with a Woodcock
// lambda. Here OptiX is gone: a workgroup of four wave64 owns a 16x16 pixel
// block, each wave an 8x8 quad (rays of a wave start close together and walk
// the same macrocells); blocks are issued in Hilbert order and spread over the
// eight XCDs so that each XCD's L2 works on a compact image region; transfer
// functions sit in LDS; voxels are read in their native dtype from planar
// channels in HBM.
#include "mdwt_device.hpp"
#include "mdwt_internal.hpp"

namespace mdwt {
namespace {

struct ChanTf {
  float dlo, dhi, xlo, xhi, os;
  int n;
  int off;  // texel offset in LDS
  int pad;
};

struct Counts {
  uint32_t samples, visits, rejections, rays;
};

struct HitRec {
  bool missed;
  float r, g, b;
  float tHit;
  int channel;
};

template <int DT>
struct Ctx {
  const uint8_t* vol;
  uint64_t chan_stride;
  int nx, ny, nz;
  float scale, offset;
  const float4* tf;      // LDS
  const ChanTf* chtf;    // LDS
  const float* maj;
  int C;
  int mx, my, mz;
  float dt;

  // sampleVolumeTexture (deviceCode.cu:122-132): NaN outside [0,1]^3, else
  // clamp-to-edge trilinear at texel coordinate p*N - 0.5.
  __device__ __forceinline__ float sample(V3 p, int c) const {
    if (p.x < 0.0f || p.x > 1.0f || p.y < 0.0f || p.y > 1.0f || p.z < 0.0f || p.z > 1.0f)
      return qnan();
    const uint8_t* base = vol + size_t(c) * chan_stride;
    const float x = p.x * float(nx) - 0.5f;
    const float y = p.y * float(ny) - 0.5f;
    const float z = p.z * float(nz) - 0.5f;
    const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
    const float fx = x - flx, fy = y - fly, fz = z - flz;
    const uint32_t x0 = uint32_t(clamp_index(flx, nx)), x1 = uint32_t(clamp_index_plus1(flx, nx));
    const uint32_t y0 = uint32_t(clamp_index(fly, ny)), y1 = uint32_t(clamp_index_plus1(fly, ny));
    const uint32_t z0 = uint32_t(clamp_index(flz, nz)), z1 = uint32_t(clamp_index_plus1(flz, nz));
    const uint32_t sx = uint32_t(nx), sxy = uint32_t(nx) * uint32_t(ny);
    const uint32_t r00 = y0 * sx + z0 * sxy, r10 = y1 * sx + z0 * sxy;
    const uint32_t r01 = y0 * sx + z1 * sxy, r11 = y1 * sx + z1 * sxy;
    const float v000 = fetch_raw<DT>(base, r00 + x0), v100 = fetch_raw<DT>(base, r00 + x1);
    const float v010 = fetch_raw<DT>(base, r10 + x0), v110 = fetch_raw<DT>(base, r10 + x1);
    const float v001 = fetch_raw<DT>(base, r01 + x0), v101 = fetch_raw<DT>(base, r01 + x1);
    const float v011 = fetch_raw<DT>(base, r11 + x0), v111 = fetch_raw<DT>(base, r11 + x1);
    const float c00 = v000 + fx * (v100 - v000);
    const float c10 = v010 + fx * (v110 - v010);
    const float c01 = v001 + fx * (v101 - v001);
    const float c11 = v011 + fx * (v111 - v011);
    const float c0 = c00 + fy * (c10 - c00);
    const float c1 = c01 + fy * (c11 - c01);
    const float r = c0 + fz * (c1 - c0);
    return r * scale + offset;
  }

  // transferFunction (deviceCode.cu:27-42), opacity only; colour is looked
  // up only for the accepted collision (same lerp, same texels).
  __device__ __forceinline__ float tf_alpha(float f, int c, float& r2out) const {
    const ChanTf t = chtf[c];
    if (f < t.dlo || f > t.dhi) {
      r2out = qnan();
      return 0.0f;
    }
    const float r1 = (f - t.dlo) / (t.dhi - t.dlo);
    const float r2 = (r1 - t.xlo) / (t.xhi - t.xlo);
    r2out = r2;
    return tf_lerp_alpha(tf + t.off, t.n, r2) * t.os;
  }
  __device__ __forceinline__ void tf_rgb(float r2, int c, HitRec& h) const {
    const ChanTf t = chtf[c];
    if (r2 != r2) {  // outside the value domain: (1, 0, 1, 0)
      h.r = 1.f;
      h.g = 0.f;
      h.b = 1.f;
      return;
    }
    const float4 v = tf_lerp(tf + t.off, t.n, r2);
    h.r = v.x;
    h.g = v.y;
    h.b = v.z;
  }
};

// -m 2, altMultiMajDTCH (deviceCode.cu:747-856)
template <int DT>
__device__ void trace_sync(const Ctx<DT>& X, const GridRay& g, float prd_t1, Rng& rng, HitRec& h,
                           Counts& cnt) {
  Dda d;
  if (!d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz)) return;
  const float step = X.dt;
  while (true) {
    cnt.visits++;
    float t0, t1;
    d.range(t0, t1);
    const int cid = d.cx + d.cy * X.mx + d.cz * X.mx * X.my;
    float minT = t1;
    bool terminate = false;
    for (int c = 0; c < X.C; ++c) {
      const float majorant = X.maj[size_t(cid) * X.C + c];
      float t = t0;
      if (majorant == 0.0f) continue;
      while (true) {
        t = t - (plogf(1.0f - rng.next()) / majorant) * step * g.worldToGridT;
        if (t >= minT) break;
        const float tWorld = t * g.gridToWorldT;
        const V3 xT = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
        if (tWorld >= prd_t1) {
          cnt.rejections++;
          if (c == X.C - 1) terminate = true;
          break;
        }
        const float thr = rng.next() * majorant;
        const float value = X.sample(xT, c);
        cnt.samples++;
        if (!(value != value)) {
          float r2;
          const float a = X.tf_alpha(value, c, r2);
          if (a > 0.0f && thr < a) {
            h.missed = false;
            h.tHit = tWorld;
            h.channel = c;
            X.tf_rgb(r2, c, h);
            minT = fminf(t, minT);
          }
        }
        cnt.rejections++;
      }
      if (terminate) break;
    }
    if (terminate || !h.missed) return;
    if (!d.advance()) return;
  }
}

// -m 0, baseLineDTCH (deviceCode.cu:858-970): one full DDA per channel; a
// channel's walk stops once it passes the nearest hit found so far.
template <int DT>
__device__ void trace_ndda(const Ctx<DT>& X, const GridRay& g, float prd_t1, Rng& rng, HitRec& h,
                           Counts& cnt) {
  Dda d0;
  if (!d0.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz)) return;
  const float step = X.dt;
  float tMaxHit = 1e20f;
  for (int c = 0; c < X.C; ++c) {
    Dda d = d0;
    while (true) {
      cnt.visits++;
      float t0, t1;
      d.range(t0, t1);
      const int cid = d.cx + d.cy * X.mx + d.cz * X.mx * X.my;
      const float majorant = X.maj[size_t(cid) * X.C + c];
      bool stop = false;
      if (majorant != 0.00f) {
        float t = t0;
        while (true) {
          t = t - (plogf(1.0f - rng.next()) / majorant) * step * g.worldToGridT;
          if (t >= t1) break;
          if (tMaxHit < t) {
            stop = true;
            break;
          }
          const float tWorld = t * g.gridToWorldT;
          const V3 xT = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
          if (tWorld >= prd_t1) {
            cnt.rejections++;
            stop = true;
            break;
          }
          const float thr = rng.next() * majorant;
          const float value = X.sample(xT, c);
          cnt.samples++;
          if (value != value) continue;
          float r2;
          const float a = X.tf_alpha(value, c, r2);
          if (a > 0.0f && thr < a) {
            h.missed = false;
            h.tHit = tWorld;
            h.channel = c;
            X.tf_rgb(r2, c, h);
            tMaxHit = fminf(tMaxHit, t);
            stop = true;
            break;
          }
          cnt.rejections++;
        }
      }
      if (stop) break;
      if (!d.advance()) break;
    }
  }
}

// -m 3, cummilativeDTCH (deviceCode.cu:513-617)
template <int DT>
__device__ void trace_cumulative(const Ctx<DT>& X, const GridRay& g, float prd_t1, Rng& rng,
                                 HitRec& h, Counts& cnt) {
  Dda d;
  if (!d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz)) return;
  const float step = X.dt;
  while (true) {
    cnt.visits++;
    float t0, t1;
    d.range(t0, t1);
    const int cid = d.cx + d.cy * X.mx + d.cz * X.mx * X.my;
    const float majorant = X.maj[cid];
    if (majorant != 0.00f) {
      float t = t0;
      while (true) {
        t = t - (plogf(1.0f - rng.next()) / majorant) * step * g.worldToGridT;
        if (t >= t1) break;
        const float tWorld = t * g.gridToWorldT;
        const V3 xT = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
        if (tWorld >= prd_t1) {
          cnt.rejections++;
          return;
        }
        float thr = rng.next() * majorant;
        for (int c = 0; c < X.C; ++c) {
          const float value = X.sample(xT, c);
          cnt.samples++;
          if (value != value) continue;
          float r2;
          const float a = X.tf_alpha(value, c, r2);
          if (a > 0.0f && thr < a) {
            h.missed = false;
            h.tHit = tWorld;
            h.channel = c;
            X.tf_rgb(r2, c, h);
            return;
          }
          thr -= a;
        }
        cnt.rejections++;
      }
    }
    if (!d.advance()) return;
  }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Workgroup -> 16x16 pixel block. Workgroups are dispatched round-robin over
// the 8 XCDs (workgroup i runs on XCD i % 8); remapping so that XCD k takes a
// contiguous run of the (Hilbert-ordered) block list keeps each XCD's L2 on a
// compact image region.
__device__ __forceinline__ uint32_t block_slot(uint32_t b, uint32_t n) {
  const uint32_t n8 = n & ~7u;
  if (b >= n8) return b;
  return (b & 7u) * (n8 >> 3) + (b >> 3);
}

template <int MODE, int DT, bool COUNT>
__global__ __launch_bounds__(256) void render_kernel(RenderParams P) {
  extern __shared__ float4 s_tf[];
  __shared__ ChanTf s_ch[kMaxChannels];
  const int C = P.vol.channels;
  // transfer functions -> LDS
  {
    const float4* g = reinterpret_cast<const float4*>(P.tf.tab);
    const int total = C * P.tf.stride;
    for (int i = threadIdx.x; i < total; i += blockDim.x) s_tf[i] = g[i];
    if (threadIdx.x < C) {
      const int c = threadIdx.x;
      ChanTf t;
      t.dlo = P.tf.dom_lo[c];
      t.dhi = P.tf.dom_hi[c];
      t.xlo = P.tf.xf_lo[c];
      t.xhi = P.tf.xf_hi[c];
      t.os = P.tf.opacity[c];
      t.n = P.tf.n[c];
      t.off = c * P.tf.stride;
      t.pad = 0;
      s_ch[c] = t;
    }
  }
  __syncthreads();

  const uint32_t slot = block_slot(blockIdx.x, uint32_t(P.fr.n_blocks));
  const uint32_t bxy = P.fr.block_order[slot];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int px = int(bxy & 0xffffu) * kBlockPixels + (wave & 1) * 8 + (lane & 7);
  const int py = int(bxy >> 16) * kBlockPixels + (wave >> 1) * 8 + (lane >> 3);
  const int W = P.fr.width, H = P.fr.height;

  Ctx<DT> X;
  X.vol = P.vol.vol;
  X.chan_stride = P.vol.chan_stride;
  X.nx = P.vol.nx;
  X.ny = P.vol.ny;
  X.nz = P.vol.nz;
  X.scale = P.vol.scale;
  X.offset = P.vol.offset;
  X.tf = s_tf;
  X.chtf = s_ch;
  X.maj = P.maj;
  X.C = C;
  X.mx = P.mc[0];
  X.my = P.mc[1];
  X.mz = P.mc[2];
  X.dt = P.dt;

  Counts cnt = {0, 0, 0, 0};
  if (px < W && py < H) {
    const uint32_t pixel = uint32_t(px) + uint32_t(W) * uint32_t(py);
    const V3 blo = load3(P.vol.blo), bhi = load3(P.vol.bhi);
    float fr = 0.f, fg = 0.f, fb = 0.f, fa = 1.f;
    for (int s = 0; s < P.fr.spp; ++s) {
      Rng rng;
      rng.init(P.fr.seed, pixel, P.fr.frame_id, uint32_t(s), 0u);
      const float jit = rng.next();
      V3 org, dir;
      camera_ray(P.cam, px, py, W, H, jit, org, dir);
      HitRec h;
      h.missed = true;
      float t0, t1;
      if (clip_root(org, dir, blo, bhi, t0, t1)) {
        const GridRay g = make_grid_ray(org, dir, blo, bhi, X.mx, X.my, X.mz);
        if (MODE == 0)
          trace_ndda<DT>(X, g, t1, rng, h, cnt);
        else if (MODE == 2)
          trace_sync<DT>(X, g, t1, rng, h, cnt);
        else
          trace_cumulative<DT>(X, g, t1, rng, h, cnt);
      }
      cnt.rays++;
      // shading with diffuse 1, specular 0, shadow 1 (deviceCode.cu:326-433)
      float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
      if (!h.missed) {
        const float li = 1.0f * 1.0f * P.fr.light_intensity;
        const float sp = 0.0f * 1.0f * P.fr.light_intensity;
        const float mr = h.r * 0.75f + P.fr.spec[0] * 0.25f;
        const float mg = h.g * 0.75f + P.fr.spec[1] * 0.25f;
        const float mb = h.b * 0.75f + P.fr.spec[2] * 0.25f;
        cr = li * h.r + sp * mr + P.fr.ambient * mr;
        cg = li * h.g + sp * mg + P.fr.ambient * mg;
        cb = li * h.b + sp * mb + P.fr.ambient * mb;
        ca = 1.0f;
      }
      // over(color, (bg, 1)) (deviceCode.cu:44-59)
      fr = fr + (cr + P.fr.bg[0] * 1.f * (1.f - ca));
      fg = fg + (cg + P.fr.bg[1] * 1.f * (1.f - ca));
      fb = fb + (cb + P.fr.bg[2] * 1.f * (1.f - ca));
      fa = fa + (ca + (1.f - ca) * 1.f);
    }
    if (!COUNT) {
      const float spp = float(P.fr.spp);
      fr = fr / spp;
      fg = fg / spp;
      fb = fb / spp;
      fa = fa / spp;
      float4* acc = reinterpret_cast<float4*>(P.fr.accum) + pixel;
      float4* sq = reinterpret_cast<float4*>(P.fr.m2) + pixel;
      const float4 q = make_float4(fr * fr, fg * fg, fb * fb, fa * fa);
      if (P.fr.accumulate) {
        // running mean (deviceCode.cu:463-469) and running mean of squares
        const float k = float(P.fr.accum_id);
        const float k1 = float(P.fr.accum_id + 1);
        const float4 a = *acc;
        const float4 m = *sq;
        *acc = make_float4((fr + k * a.x) / k1, (fg + k * a.y) / k1, (fb + k * a.z) / k1,
                           (fa + k * a.w) / k1);
        *sq = make_float4((q.x + k * m.x) / k1, (q.y + k * m.y) / k1, (q.z + k * m.z) / k1,
                          (q.w + k * m.w) / k1);
      } else {
        *acc = make_float4(fr, fg, fb, fa);
        *sq = q;
      }
    }
  }
  if (COUNT) {
    const uint32_t s = wave_sum(cnt.samples), v = wave_sum(cnt.visits);
    const uint32_t r = wave_sum(cnt.rejections), n = wave_sum(cnt.rays);
    if (lane == 0) {
      atomicAdd(&P.fr.counters[0], (unsigned long long)s);
      atomicAdd(&P.fr.counters[1], (unsigned long long)v);
      atomicAdd(&P.fr.counters[2], (unsigned long long)r);
      atomicAdd(&P.fr.counters[3], (unsigned long long)n);
    }
  }
}

// ---------------------------------------------------------- test hooks -----
__global__ void gen_rays_kernel(CameraParams cam, int W, int H, uint32_t seed, uint32_t frame,
                                uint32_t sample, float* org, float* dir) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int px = i % W, py = i / W;
  Rng rng;
  rng.init(seed, uint32_t(i), frame, sample, 0u);
  const float jit = rng.next();
  V3 o, d;
  camera_ray(cam, px, py, W, H, jit, o, d);
  org[3 * i] = o.x;
  org[3 * i + 1] = o.y;
  org[3 * i + 2] = o.z;
  dir[3 * i] = d.x;
  dir[3 * i + 1] = d.y;
  dir[3 * i + 2] = d.z;
}

__global__ void clip_kernel(VolumeParams vol, const float* org, const float* dir, int n, float* t01,
                            int* hit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float t0 = 0.f, t1 = 0.f;
  const bool h = clip_root(load3(org + 3 * i), load3(dir + 3 * i), load3(vol.blo), load3(vol.bhi),
                           t0, t1);
  t01[2 * i] = t0;
  t01[2 * i + 1] = t1;
  hit[i] = h ? 1 : 0;
}

__global__ void trace_cells_kernel(VolumeParams vol, int mx, int my, int mz, const float* org,
                                   const float* dir, int n, int max_cells, int* cells, float* t01,
                                   int* counts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GridRay g = make_grid_ray(load3(org + 3 * i), load3(dir + 3 * i), load3(vol.blo),
                                  load3(vol.bhi), mx, my, mz);
  Dda d;
  int k = 0;
  if (d.begin(g.org, g.dir, 1e20f, mx, my, mz)) {
    while (true) {
      float t0, t1;
      d.range(t0, t1);
      if (k < max_cells) {
        const size_t o = size_t(i) * max_cells + k;
        cells[o] = d.cx + d.cy * mx + d.cz * mx * my;
        t01[2 * o] = t0;
        t01[2 * o + 1] = t1;
      }
      ++k;
      if (!d.advance()) break;
    }
  }
  for (int j = k; j < max_cells; ++j) {  // unused slots: cell -1, t = 0
    const size_t o = size_t(i) * max_cells + j;
    cells[o] = -1;
    t01[2 * o] = 0.f;
    t01[2 * o + 1] = 0.f;
  }
  counts[i] = k;
}

// make_rgba (OWL) of the accumulated colour: min(255, max(0, int(256 c)))
__device__ __forceinline__ uint32_t to8(float f) {
  const float c = fminf(fmaxf(f * 256.f, -1.f), 256.f);
  const int i = int(c);
  return uint32_t(i < 0 ? 0 : (i > 255 ? 255 : i));
}
__global__ void rgba8_kernel(const float4* acc, uint32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 c = acc[i];
  out[i] = to8(c.x) | (to8(c.y) << 8) | (to8(c.z) << 16) | (to8(c.w) << 24);
}

__global__ __launch_bounds__(256) void copy_kernel(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

template <int MODE, bool COUNT>
hipError_t launch_render_dt(const RenderParams& p, int dtype, size_t lds, hipStream_t s) {
  const dim3 grid(unsigned(p.fr.n_blocks)), block(kThreads);
  switch (dtype) {
    case MDWT_DT_U8:
      hipLaunchKernelGGL((render_kernel<MODE, MDWT_DT_U8, COUNT>), grid, block, lds, s, p);
      break;
    case MDWT_DT_U16:
      hipLaunchKernelGGL((render_kernel<MODE, MDWT_DT_U16, COUNT>), grid, block, lds, s, p);
      break;
    case MDWT_DT_F16:
      hipLaunchKernelGGL((render_kernel<MODE, MDWT_DT_F16, COUNT>), grid, block, lds, s, p);
      break;
    default:
      hipLaunchKernelGGL((render_kernel<MODE, MDWT_DT_F32, COUNT>), grid, block, lds, s, p);
      break;
  }
  return hipGetLastError();
}

template <int MODE, bool COUNT>
hipError_t set_lds_limit_dt(size_t lds) {
  hipError_t e = hipSuccess;
  const int v = int(lds);
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(render_kernel<MODE, MDWT_DT_U8, COUNT>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, v)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(render_kernel<MODE, MDWT_DT_U16, COUNT>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, v)) != hipSuccess)
    return e;
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(render_kernel<MODE, MDWT_DT_F16, COUNT>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, v)) != hipSuccess)
    return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(render_kernel<MODE, MDWT_DT_F32, COUNT>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, v);
}

}  // namespace

hipError_t launch_render(const RenderParams& p, int mode, int dtype, bool count, hipStream_t s) {
  const size_t lds = size_t(p.vol.channels) * size_t(p.tf.stride) * 16;
  if (lds > 65536) {
    hipError_t e = hipSuccess;
    if (mode == 0) e = count ? set_lds_limit_dt<0, true>(lds) : set_lds_limit_dt<0, false>(lds);
    else if (mode == 2) e = count ? set_lds_limit_dt<2, true>(lds) : set_lds_limit_dt<2, false>(lds);
    else e = count ? set_lds_limit_dt<3, true>(lds) : set_lds_limit_dt<3, false>(lds);
    if (e != hipSuccess) return e;
  }
  if (mode == 0) return count ? launch_render_dt<0, true>(p, dtype, lds, s)
                              : launch_render_dt<0, false>(p, dtype, lds, s);
  if (mode == 2) return count ? launch_render_dt<2, true>(p, dtype, lds, s)
                              : launch_render_dt<2, false>(p, dtype, lds, s);
  return count ? launch_render_dt<3, true>(p, dtype, lds, s)
               : launch_render_dt<3, false>(p, dtype, lds, s);
}

hipError_t launch_gen_rays(const CameraParams& cam, int W, int H, uint32_t seed, uint32_t frame,
                           uint32_t sample, float* org, float* dir, hipStream_t s) {
  const int n = W * H;
  hipLaunchKernelGGL(gen_rays_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, cam, W, H,
                     seed, frame, sample, org, dir);
  return hipGetLastError();
}

hipError_t launch_clip(const VolumeParams& vol, const float* org, const float* dir, int n,
                       float* t01, int* hit, hipStream_t s) {
  hipLaunchKernelGGL(clip_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, vol, org, dir,
                     n, t01, hit);
  return hipGetLastError();
}

hipError_t launch_trace_cells(const VolumeParams& vol, const int mc[3], const float* org,
                              const float* dir, int n, int max_cells, int* cells, float* t01,
                              int* counts, hipStream_t s) {
  hipLaunchKernelGGL(trace_cells_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, vol,
                     mc[0], mc[1], mc[2], org, dir, n, max_cells, cells, t01, counts);
  return hipGetLastError();
}

hipError_t launch_rgba8(const void* accum, uint32_t* out, int n, hipStream_t s) {
  hipLaunchKernelGGL(rgba8_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(accum), out, n);
  return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t n16, hipStream_t s) {
  hipLaunchKernelGGL(copy_kernel, dim3(2048u), dim3(256), 0, s, reinterpret_cast<const float4*>(src),
                     reinterpret_cast<float4*>(dst), n16);
  return hipGetLastError();
}

}  // namespace mdwt
