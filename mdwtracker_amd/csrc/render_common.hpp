// Shared device pieces of the frame kernels: per-channel transfer functions
// in LDS, volume sampling, work counters and the block schedule.
#pragma once
#include "mdwt_device.hpp"

namespace mdwt {

// Per-channel transfer-function parameters in LDS: value domain [dlo, dhi],
// TF range start xlo, the correctly rounded reciprocals of the domain width
// (idw) and of the TF range width (ixw), opacity scale, texel count and
// texel offset.
struct ChanTf {
  float dlo, dhi, xlo, ixw, os;
  int n;
  int off;  // texel offset in LDS
  float idw;
};

struct Counts {
  uint32_t samples, visits, rejections, rays;
};

struct HitRec {
  bool missed;
  float r, g, b;
  float a;  // transparency: 1 for a Woodcock hit, the composited alpha for the marchers
  float tHit;
  int channel;
};

template <int DT>
struct Ctx {
  const uint8_t* bricks;
  uint64_t brick_stride;
  uint32_t nbx, nbxy;  // bricks per row and per slab (both below 2^24)
  int nx, ny, nz;
  float scale, offset;
  const float4* tf;    // LDS
  const ChanTf* chtf;  // LDS
  const float* maj;
  int C;
  int mx, my, mz;
  float dt;

  __device__ __forceinline__ void init(const RenderParams& P, const float4* tf_lds,
                                       const ChanTf* ch_lds) {
    bricks = P.vol.bricks;
    brick_stride = P.vol.brick_stride;
    nbx = uint32_t(P.vol.nbx);
    nbxy = nbx * uint32_t(P.vol.nby);
    nx = P.vol.nx;
    ny = P.vol.ny;
    nz = P.vol.nz;
    scale = P.vol.scale;
    offset = P.vol.offset;
    tf = tf_lds;
    chtf = ch_lds;
    maj = P.maj;
    C = P.vol.channels;
    mx = P.mc[0];
    my = P.mc[1];
    mz = P.mc[2];
    dt = P.dt;
  }

  // sampleVolumeTexture (deviceCode.cu:122-132): NaN outside [0,1]^3, else
  // clamp-to-edge trilinear at texel coordinate p*N - 0.5.
  __device__ __forceinline__ float sample(V3 p, int c) const {
    if (p.x < 0.0f || p.x > 1.0f || p.y < 0.0f || p.y > 1.0f || p.z < 0.0f || p.z > 1.0f)
      return qnan();
    const float x = p.x * float(nx) - 0.5f;
    const float y = p.y * float(ny) - 0.5f;
    const float z = p.z * float(nz) - 0.5f;
    const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
    const float fx = x - flx, fy = y - fly, fz = z - flz;
    const uint32_t x0 = uint32_t(clamp_index(flx, nx)), x1 = uint32_t(clamp_index_plus1(flx, nx));
    const uint32_t y0 = uint32_t(clamp_index(fly, ny)), y1 = uint32_t(clamp_index_plus1(fly, ny));
    const uint32_t z0 = uint32_t(clamp_index(flz, nz)), z1 = uint32_t(clamp_index_plus1(flz, nz));
    // the voxels are read from the 4 x 4 x 4 brick copy: index = X(x) +
    // Y(y) + Z(z) (a trilinear footprint touches about 1.8 cache lines
    // instead of 4 in the planar layout; 3.5% faster at -m 2 on 4 x 256^3 u8)
    const uint8_t* base = bricks + size_t(c) * brick_stride;
    const uint32_t ax0 = ((x0 >> 2) << 6) | (x0 & 3u), ax1 = ((x1 >> 2) << 6) | (x1 & 3u);
    // 24-bit products (full-rate v_mul_u32_u24): brick coordinates and the
    // bricks per row / slab are below 2^24 (checked when the volume is set)
    const uint32_t ay0 = (__umul24(y0 >> 2, nbx) << 6) | ((y0 & 3u) << 2);
    const uint32_t ay1 = (__umul24(y1 >> 2, nbx) << 6) | ((y1 & 3u) << 2);
    const uint32_t az0 = (__umul24(z0 >> 2, nbxy) << 6) | ((z0 & 3u) << 4);
    const uint32_t az1 = (__umul24(z1 >> 2, nbxy) << 6) | ((z1 & 3u) << 4);
    const uint32_t r00 = ay0 + az0, r10 = ay1 + az0, r01 = ay0 + az1, r11 = ay1 + az1;
    const float v000 = fetch_raw<DT>(base, r00 + ax0), v100 = fetch_raw<DT>(base, r00 + ax1);
    const float v010 = fetch_raw<DT>(base, r10 + ax0), v110 = fetch_raw<DT>(base, r10 + ax1);
    const float v001 = fetch_raw<DT>(base, r01 + ax0), v101 = fetch_raw<DT>(base, r01 + ax1);
    const float v011 = fetch_raw<DT>(base, r11 + ax0), v111 = fetch_raw<DT>(base, r11 + ax1);
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
  // up only for the accepted collision (same lerp, same texels). The two
  // divisions by per-channel constants are products with their reciprocals.
  __device__ __forceinline__ float tf_alpha(float f, int c, float& r2out) const {
    const ChanTf t = chtf[c];
    if (f < t.dlo || f > t.dhi) {
      r2out = qnan();
      return 0.0f;
    }
    const float r1 = (f - t.dlo) * t.idw;
    const float r2 = (r1 - t.xlo) * t.ixw;
    r2out = r2;
    return tf_lerp_alpha(tf + t.off, t.n, r2) * t.os;
  }
  // full transferFunction (deviceCode.cu:27-42) for the blend and marcher
  // modes; f is never NaN here (callers test the unit cube or the sample)
  __device__ __forceinline__ float4 tf_rgba(float f, int c) const {
    const ChanTf t = chtf[c];
    if (f < t.dlo || f > t.dhi) return make_float4(1.f, 0.f, 1.f, 0.f);
    const float r1 = (f - t.dlo) * t.idw;
    const float r2 = (r1 - t.xlo) * t.ixw;
    float4 v = tf_lerp(tf + t.off, t.n, r2);
    v.w = v.w * t.os;
    return v;
  }
  // calculateGradient (deviceCode.cu:243-281): central differences of the
  // channel's TF opacity, one-sided at the faces of the unit cube
  __device__ V3 gradient(V3 p, float eps, int c) const {
    const float pc[3] = {p.x, p.y, p.z};
    float g[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      float p1[3] = {pc[0], pc[1], pc[2]}, p2[3] = {pc[0], pc[1], pc[2]};
      p1[i] -= eps;
      p2[i] += eps;
      float div;
      if (p1[i] < 0.0f) {
        p1[i] = pc[i];
        p2[i] = pc[i] + eps;
        div = eps;
      } else if (p2[i] > 1.0f) {
        p1[i] = pc[i] - eps;
        p2[i] = pc[i];
        div = eps;
      } else {
        div = 2.0f * eps;
      }
      float r;
      const float v1 = tf_alpha(sample(v3(p1[0], p1[1], p1[2]), c), c, r);
      const float v2 = tf_alpha(sample(v3(p2[0], p2[1], p2[2]), c), c, r);
      g[i] = (v2 - v1) / div;
    }
    return v3(g[0], g[1], g[2]);
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

// Transfer-function texels and per-channel TF parameters -> LDS (the caller
// synchronises the workgroup afterwards).
__device__ __forceinline__ void load_tf_lds(const RenderParams& P, float4* s_tf, ChanTf* s_ch) {
  const int C = P.vol.channels;
  const float4* g = reinterpret_cast<const float4*>(P.tf.tab);
  const int total = C * P.tf.stride;
  for (int i = threadIdx.x; i < total; i += blockDim.x) s_tf[i] = g[i];
  if (int(threadIdx.x) < C) {
    const int c = threadIdx.x;
    ChanTf t;
    t.dlo = P.tf.dom_lo[c];
    t.dhi = P.tf.dom_hi[c];
    t.xlo = P.tf.xf_lo[c];
    t.ixw = 1.0f / (P.tf.xf_hi[c] - P.tf.xf_lo[c]);
    t.os = P.tf.opacity[c];
    t.n = P.tf.n[c];
    t.off = c * P.tf.stride;
    t.idw = 1.0f / (P.tf.dom_hi[c] - P.tf.dom_lo[c]);
    s_ch[c] = t;
  }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void flush_counts(const RenderParams& P, const Counts& cnt) {
  const uint32_t s = wave_sum(cnt.samples), v = wave_sum(cnt.visits);
  const uint32_t r = wave_sum(cnt.rejections), n = wave_sum(cnt.rays);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&P.fr.counters[0], (unsigned long long)s);
    atomicAdd(&P.fr.counters[1], (unsigned long long)v);
    atomicAdd(&P.fr.counters[2], (unsigned long long)r);
    atomicAdd(&P.fr.counters[3], (unsigned long long)n);
  }
}

// Workgroup -> 16x16 pixel block (index into the Hilbert block order).
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i % 8), so the identity map deals consecutive Hilbert blocks to
// different XCDs: every XCD gets an even share of the heavy image regions
// (giving each XCD a contiguous run instead left half of them idle for half
// the frame: the volume's footprint is not spread evenly along the curve),
// and each XCD still works along the curve, every 8th block. Dealing runs of
// 2, 4, 8 or 16 consecutive blocks to each XCD in turn (more L2 sharing
// between neighbouring blocks) measured 0.5-2% slower on 4 x 256^3 u8 at
// 1024^2.
__device__ __forceinline__ uint32_t block_slot(uint32_t b, uint32_t /*n*/) { return b; }

// running mean of the colour (deviceCode.cu:463-469) and, when m2 is set, of
// its square (an addition for confidence intervals and the error estimate;
// off by default, so a plain frame moves only the accumulator's 32 bytes per
// pixel)
__device__ __forceinline__ void accumulate_pixel(const RenderParams& P, void* accum, void* m2,
                                                 uint32_t accum_id, uint32_t pixel, float fr,
                                                 float fg, float fb, float fa) {
  float4* acc = reinterpret_cast<float4*>(accum) + pixel;
  if (P.fr.accumulate) {
    const float k = float(accum_id);
    const float k1 = float(accum_id + 1);
    const float4 a = *acc;
    *acc = make_float4((fr + k * a.x) / k1, (fg + k * a.y) / k1, (fb + k * a.z) / k1,
                       (fa + k * a.w) / k1);
  } else {
    *acc = make_float4(fr, fg, fb, fa);
  }
  if (m2) {
    float4* sq = reinterpret_cast<float4*>(m2) + pixel;
    const float4 q = make_float4(fr * fr, fg * fg, fb * fb, fa * fa);
    if (P.fr.accumulate) {
      const float k = float(accum_id);
      const float k1 = float(accum_id + 1);
      const float4 m = *sq;
      *sq = make_float4((q.x + k * m.x) / k1, (q.y + k * m.y) / k1, (q.z + k * m.z) / k1,
                        (q.w + k * m.w) / k1);
    } else {
      *sq = q;
    }
  }
}

// Shading of one sample (deviceCode.cu:411-433): diffuse, specular and shadow
// factors from the caller (1, 0, 1 without gradient shading and shadows),
// the hit channel's material specular colour, then over the background.
__device__ __forceinline__ void composite_sample(const RenderParams& P, bool hit, float hr,
                                                 float hg, float hb, float ha, float diffuse,
                                                 float specular, float shadow, int ch, float& fr,
                                                 float& fg, float& fb, float& fa) {
  float cr = 0.f, cg = 0.f, cb = 0.f, ca = 0.f;
  if (hit) {
    const float ld = diffuse * shadow * P.fr.light_intensity;
    const float ls = specular * shadow * P.fr.light_intensity;
    const float* sc = P.fr.sh.spec[ch];
    const float mr = hr * 0.75f + sc[0] * 0.25f;
    const float mg = hg * 0.75f + sc[1] * 0.25f;
    const float mb = hb * 0.75f + sc[2] * 0.25f;
    cr = ld * hr + ls * mr + P.fr.ambient * mr;
    cg = ld * hg + ls * mg + P.fr.ambient * mg;
    cb = ld * hb + ls * mb + P.fr.ambient * mb;
    ca = ha;
  }
  // over(color, (bg, 1)) (deviceCode.cu:54-59, 433)
  fr = fr + (cr + P.fr.bg[0] * 1.f * (1.f - ca));
  fg = fg + (cg + P.fr.bg[1] * 1.f * (1.f - ca));
  fb = fb + (cb + P.fr.bg[2] * 1.f * (1.f - ca));
  fa = fa + (ca + (1.f - ca) * 1.f);
}

// Divides the pixel sums (alpha started at 1) by spp and accumulates.
// What one pixel block needs from its view: camera, size, samples, seed,
// frame / accumulation ids and accumulators.
struct ViewCtx {
  const CameraParams* cam;
  int W, H, spp;
  uint32_t seed, frame_id, accum_id;
  void* accum;
  void* m2;
};

__device__ __forceinline__ void finish_pixel(const RenderParams& P, void* accum, void* m2,
                                             uint32_t accum_id, int spp_i, uint32_t pixel,
                                             float fr, float fg, float fb, float fa) {
  const float spp = float(spp_i);
  accumulate_pixel(P, accum, m2, accum_id, pixel, fr / spp, fg / spp, fb / spp, fa / spp);
}

}  // namespace mdwt
