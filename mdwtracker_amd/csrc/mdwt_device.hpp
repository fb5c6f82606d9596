// Device-side building blocks of the MDWT kernels (gfx950, wave64).
//
// Every fp32 expression is written in the order of the reference expression
// it restates and the library is built with -ffp-contract=off and correctly
// rounded division/sqrt, so each operation rounds once, exactly as on the
// host. That is what lets range grids, majorants, DDA cell sequences and
// (with the counter-based RNG below) whole images match the CPU oracle bit
// for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdwt_types.hpp"

namespace mdwt {

// ------------------------------------------------------------ vectors ------
struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 normalize3(V3 a) { return a * (1.f / sqrtf(dot3(a, a))); }
__device__ __forceinline__ float min3(V3 a) { return fminf(fminf(a.x, a.y), a.z); }
__device__ __forceinline__ float max3(V3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
__device__ __forceinline__ V3 load3(const float* p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ float qnan() { return __int_as_float(0x7fc00000); }

// ------------------------------------------------------------- Philox ------
// Philox4x32-7 (Salmon et al., SC'11: seven rounds is the fewest for which
// Philox4x32 passes the BigCrush battery; the round function and key schedule
// are those of the usual ten-round generator). Stream key = (seed, pixel);
// counter = (frame, sample, block, stream). Each round's two 32x32->64 products are
// written as 64-bit products so that each becomes one v_mad_u64_u32 (32-bit
// integer multiplies are quarter rate on CDNA; the hi/lo pair would be two).
constexpr int kPhiloxRounds = 7;
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < kPhiloxRounds; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = uint32_t(p1);
    c2 = n2;
    c3 = uint32_t(p0);
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// uniform in [0,1) with 24 bits: (w >> 8) * 2^-24, exact
__device__ __forceinline__ float to_unit(uint32_t w) { return float(w >> 8) * 0x1p-24f; }

// Word layout of a (pixel, frame, sample, stream) stream: the jitter is word 0
// of block (frame, sample, 0, stream), and with two-axis jitter the y jitter
// is word 1 of it.
// -m 0 / -m 1 / -m 2 track every (macrocell visit v, channel c) from the
// cell's entry (v counts the cells of the ray's DDA walk; -m 1 runs the
// channels' clocks on these positions); -m 3 tracks its summed majorant per
// visit at the positions of channel 0:
//   * the first free-flight distance of channel c is word c mod 4 of the cell
//     block (frame, sample, 1 + 16 v + 12 + c / 4, stream): one Philox block
//     gives the first step of up to four channels, and most (v, c) steps
//     leave the cell at that first step;
//   * after step j lands inside the cell, pair j of the (v, c) stream holds
//     its threshold and the next step's distance: words 2 (j mod 2) and
//     2 (j mod 2) + 1 of block (frame, sample, 1 + 16 v + c, stream |
//     2 floor(j / 2)).
// Every (v, c) step is thus a function of (v, c, j) alone, whichever lane
// or loop computes it. The other modes (4, 5) count steps from the ray's start:
// step j owns words 2 (j mod 2) (distance) and 2 (j mod 2) + 1 (threshold)
// of block (frame, sample, 1, stream | 2 floor(j / 2)).
struct Rng {
  uint32_t k0, k1, frame, sample, stream;
  uint32_t k, cur, base;
  uint32_t b0, b1, b2, b3;
  __device__ __forceinline__ void seek(uint32_t blk) {
    base = blk;
    k = 0;
  }
  __device__ __forceinline__ void init(uint32_t seed, uint32_t pixel, uint32_t frame_, uint32_t sample_,
                                       uint32_t stream_) {
    k0 = seed;
    k1 = pixel;
    frame = frame_;
    sample = sample_;
    stream = stream_;
    k = 0;
    cur = 0;
    base = 0;
  }
  // the pixel jitter of x and y: word 0 of the jitter block for both axes
  // (deviceCode.cu:298), or words 0 and 1 with two_d
  __device__ __forceinline__ void jitter2(bool two_d, float& jx, float& jy) const {
    uint32_t o[4];
    philox(frame, sample, 0u, stream, k0, k1, o);
    jx = to_unit(o[0]);
    jy = two_d ? to_unit(o[1]) : jx;
  }
  __device__ __forceinline__ float step() {
    cur = k;
    k += 1;
    if ((cur & 1u) == 0u) {
      uint32_t o[4];
      philox(frame, sample, 1u + base, stream | ((cur >> 1) << 1), k0, k1, o);
      b0 = o[0];
      b1 = o[1];
      b2 = o[2];
      b3 = o[3];
    }
    return to_unit((cur & 1u) ? b2 : b0);
  }
  __device__ __forceinline__ float thr() const { return to_unit((cur & 1u) ? b3 : b1); }

  // -m 0 / -m 2: the cell block of channel group g at visit v
  __device__ __forceinline__ void cell_block(uint32_t v, uint32_t g, uint32_t o[4]) const {
    philox(frame, sample, 1u + ((v << 4) | (12u + g)), stream, k0, k1, o);
  }
  // start the pairs of (visit v, channel c)
  __device__ __forceinline__ void start_pairs(uint32_t v, uint32_t c) {
    base = (v << 4) | c;
    k = 0;
  }
  // threshold of the current step (pair k, word 0); computes the pair's block
  __device__ __forceinline__ float pair_thr() {
    if ((k & 1u) == 0u) {
      uint32_t o[4];
      philox(frame, sample, 1u + base, stream | ((k >> 1) << 1), k0, k1, o);
      b0 = o[0];
      b1 = o[1];
      b2 = o[2];
      b3 = o[3];
    }
    return to_unit((k & 1u) ? b2 : b0);
  }
  // distance of the next step (pair k, word 1); moves on to the next pair
  __device__ __forceinline__ float pair_dist() {
    const uint32_t p = k++;
    return to_unit((p & 1u) ? b3 : b1);
  }
  // -m 1: pair j of the (visit v, channel c) stream, without pair state
  __device__ __forceinline__ void vc_pair(uint32_t v, uint32_t c, uint32_t j, float& th,
                                          float& dist) const {
    uint32_t o[4];
    philox(frame, sample, 1u + ((v << 4) | c), stream | ((j >> 1) << 1), k0, k1, o);
    th = to_unit((j & 1u) ? o[2] : o[0]);
    dist = to_unit((j & 1u) ? o[3] : o[1]);
  }
};

__device__ __forceinline__ uint32_t pick_word(const uint32_t w[4], uint32_t i) {
  return i == 0u ? w[0] : (i == 1u ? w[1] : (i == 2u ? w[2] : w[3]));
}

// ------------------------------------------------------------ log ----------
// fp32 natural log for normal x > 0: Cody-Waite reduction x = 2^e m,
// m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) by its series to the
// s^9 term (Abramowitz & Stegun 4.1.27). Same operations as the host side.
__device__ __forceinline__ float plogf(float x) {
  const uint32_t bits = __float_as_uint(x);
  int e = int((bits >> 23) & 0xffu) - 127;
  float m = __uint_as_float((bits & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  const float p = ((((0.11111111f * z + 0.14285715f) * z + 0.2f) * z + 0.33333334f) * z);
  const float r = (s + s) + (s + s) * p;
  const float fe = float(e);
  return fe * 0.693145751953125f + (r + fe * 1.42860677e-06f);
}

// The same reduction with log m = f P(f), f = m - 1, P of degree 9 in Horner
// form with fused multiply-adds (no division): every free-flight step uses it.
// Same operations as the host side (fmaf rounds once on both).
__device__ __forceinline__ float plogf_poly(float x) {
  const uint32_t bits = __float_as_uint(x);
  int e = int((bits >> 23) & 0xffu) - 127;
  float m = __uint_as_float((bits & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  const float f = m - 1.0f;
  float p = -7.637512684e-02f;
  p = fmaf(p, f, 1.291354448e-01f);
  p = fmaf(p, f, -1.324101686e-01f);
  p = fmaf(p, f, 1.418020874e-01f);
  p = fmaf(p, f, -1.660911888e-01f);
  p = fmaf(p, f, 2.000206262e-01f);
  p = fmaf(p, f, -2.500158548e-01f);
  p = fmaf(p, f, 3.333332539e-01f);
  p = fmaf(p, f, -4.999998808e-01f);
  p = fmaf(p, f, 1.000000000e+00f);
  const float r = f * p;
  const float fe = float(e);
  return fe * 0.693145751953125f + (r + fe * 1.42860677e-06f);
}

// fp32 exp: x = n ln2 + r (Cody-Waite, the ln2 split of plogf), e^r by its
// Taylor polynomial to r^6 (|r| <= ln2/2), times 2^n from exponent bits.
// 0 below -87.3 (no denormals), +inf above 88. Same operations as the host.
__device__ __forceinline__ float pexpf(float x) {
  if (!(x > -87.3f)) return 0.0f;
  if (x > 88.0f) return __int_as_float(0x7f800000);
  const float n = floorf(x * 1.44269504f + 0.5f);
  const float r = (x - n * 0.693145751953125f) - n * 1.42860677e-06f;
  const float p =
      1.0f + r * (1.0f + r * (0.5f + r * (0.16666667f + r * (0.041666668f + r * (0.008333334f +
                                                                           r * 0.0013888889f)))));
  const int ni = int(n);
  return p * __uint_as_float(uint32_t(ni + 127) << 23);
}
// pow(x, y) for x > 0 (Blinn-Phong lobe) as exp(y log x)
__device__ __forceinline__ float ppowf(float x, float y) { return pexpf(y * plogf(x)); }

// --------------------------------------------------------- index clamps ----
__device__ __forceinline__ int clamp_index(float fl, int n) {
  const float c = fminf(fmaxf(fl, -1.0f), float(n));
  const int i = int(c);
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}
__device__ __forceinline__ int clamp_index_plus1(float fl, int n) {
  const float c = fminf(fmaxf(fl, -1.0f), float(n));
  const int i = int(c) + 1;
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

// Linear clamp-to-edge lookup of a 1-D float4 table at normalized u
// (texel coordinate u*N - 0.5, lerp a + f*(b - a)).
__device__ __forceinline__ float4 tf_lerp(const float4* tab, int n, float u) {
  const float x = u * float(n) - 0.5f;
  const float fl = floorf(x);
  const float f = x - fl;
  const float4 a = tab[clamp_index(fl, n)];
  const float4 b = tab[clamp_index_plus1(fl, n)];
  return make_float4(a.x + f * (b.x - a.x), a.y + f * (b.y - a.y), a.z + f * (b.z - a.z),
                     a.w + f * (b.w - a.w));
}
__device__ __forceinline__ float tf_lerp_alpha(const float4* tab, int n, float u) {
  const float x = u * float(n) - 0.5f;
  const float fl = floorf(x);
  const float f = x - fl;
  const float a = tab[clamp_index(fl, n)].w;
  const float b = tab[clamp_index_plus1(fl, n)].w;
  return a + f * (b - a);
}

// --------------------------------------------------------- voxel fetch -----
template <int DT>
__device__ __forceinline__ float fetch_raw(const uint8_t* base, uint32_t i);
template <>
__device__ __forceinline__ float fetch_raw<MDWT_DT_U8>(const uint8_t* base, uint32_t i) {
  return float(base[i]);
}
template <>
__device__ __forceinline__ float fetch_raw<MDWT_DT_U16>(const uint8_t* base, uint32_t i) {
  return float(reinterpret_cast<const uint16_t*>(base)[i]);
}
template <>
__device__ __forceinline__ float fetch_raw<MDWT_DT_F16>(const uint8_t* base, uint32_t i) {
  return float(reinterpret_cast<const _Float16*>(base)[i]);
}
template <>
__device__ __forceinline__ float fetch_raw<MDWT_DT_F32>(const uint8_t* base, uint32_t i) {
  return reinterpret_cast<const float*>(base)[i];
}

// ------------------------------------------------------------- DDA ---------
// Amanatides-Woo traversal of the majorant grid as an iterator: begin() sets
// up the first cell, range() gives the current cell's [t0, t1], advance()
// steps every axis whose boundary is closest (ties step several axes) and
// returns false when the ray leaves the grid. Same arithmetic as
// dda.h:82-203, including its zero-direction handling (such rays are
// dropped).
struct Dda {
  float tnx, tny, tnz;     // t_next
  float tsx, tsy, tsz;     // t_step
  int cx, cy, cz;          // cell
  int sx, sy, sz;          // stop
  int dx, dy, dz;          // cell delta
  float ray_t0, next_begin, tmax;

  __device__ __forceinline__ bool begin(V3 org, V3 dir, float tMax, int gx, int gy, int gz) {
    const float fsx = float(gx), fsy = float(gy), fsz = float(gz);
    const float rx = 1.f / dir.x, ry = 1.f / dir.y, rz = 1.f / dir.z;
    const float tlx = (0.f - org.x) * rx, tly = (0.f - org.y) * ry, tlz = (0.f - org.z) * rz;
    const float thx = (fsx - org.x) * rx, thy = (fsy - org.y) * ry, thz = (fsz - org.z) * rz;
    float nrx = fminf(tlx, thx), nry = fminf(tly, thy), nrz = fminf(tlz, thz);
    const float frx = fmaxf(tlx, thx), fry = fmaxf(tly, thy), frz = fmaxf(tlz, thz);
    if (dir.x == 0.f) {
      if (org.x < 0.f || org.x > fsx) return false;
      nrx = __int_as_float(0x7f800000);
    }
    if (dir.y == 0.f) {
      if (org.y < 0.f || org.y > fsy) return false;
      nry = __int_as_float(0x7f800000);
    }
    if (dir.z == 0.f) {
      if (org.z < 0.f || org.z > fsz) return false;
      nrz = __int_as_float(0x7f800000);
    }
    ray_t0 = fmaxf(0.f, fmaxf(fmaxf(nrx, nry), nrz));
    const float ray_t1 = fminf(tMax, fminf(fminf(frx, fry), frz));
    if (ray_t0 > ray_t1) return false;
    const float ox = org.x + ray_t0 * dir.x;
    const float oy = org.y + ray_t0 * dir.y;
    const float oz = org.z + ray_t0 * dir.z;
    const float fcx = fmaxf(0.f, fminf(fsx - 1.f, floorf(ox)));
    const float fcy = fmaxf(0.f, fminf(fsy - 1.f, floorf(oy)));
    const float fcz = fmaxf(0.f, fminf(fsz - 1.f, floorf(oz)));
    const float ex = dir.x > 0.f ? fcx + 1.f : fcx;
    const float ey = dir.y > 0.f ? fcy + 1.f : fcy;
    const float ez = dir.z > 0.f ? fcz + 1.f : fcz;
    tsx = fabsf(rx);
    tsy = fabsf(ry);
    tsz = fabsf(rz);
    const float inf = __int_as_float(0x7f800000);
    tnx = (dir.x == 0.f) ? inf : fabsf(ex - ox) * tsx;
    tny = (dir.y == 0.f) ? inf : fabsf(ey - oy) * tsy;
    tnz = (dir.z == 0.f) ? inf : fabsf(ez - oz) * tsz;
    sx = dir.x > 0.f ? gx : -1;
    sy = dir.y > 0.f ? gy : -1;
    sz = dir.z > 0.f ? gz : -1;
    dx = dir.x > 0.f ? 1 : -1;
    dy = dir.y > 0.f ? 1 : -1;
    dz = dir.z > 0.f ? 1 : -1;
    cx = int(fcx);
    cy = int(fcy);
    cz = int(fcz);
    next_begin = 0.f;
    tmax = tMax;
    return true;
  }
  __device__ __forceinline__ void range(float& t0, float& t1) const {
    const float tc = fminf(fminf(tnx, tny), tnz);
    t0 = ray_t0 + next_begin;
    t1 = ray_t0 + fminf(tc, tmax);
  }
  // For finite rays at least one axis equals the minimum; a NaN ray steps no
  // axis and is ended here instead of looping (the reference would spin).
  __device__ __forceinline__ bool advance() {
    const float tc = fminf(fminf(tnx, tny), tnz);
    next_begin = tc;
    bool stepped = false;
    if (tnx == tc) {
      tnx += tsx;
      cx += dx;
      stepped = true;
      if (cx == sx) return false;
    }
    if (tny == tc) {
      tny += tsy;
      cy += dy;
      stepped = true;
      if (cy == sy) return false;
    }
    if (tnz == tc) {
      tnz += tsz;
      cz += dz;
      stepped = true;
      if (cz == sz) return false;
    }
    return stepped;
  }
};

// World ray -> majorant-grid ray (the closest-hit programs' set-up,
// deviceCode.cu:866-888).
struct GridRay {
  V3 worldOrg, worldDir, worldToUnit, org, dir;
  float worldToGridT, gridToWorldT;
};
__device__ __forceinline__ GridRay make_grid_ray(V3 worldOrg, V3 worldDir, V3 blo, V3 bhi, int mx,
                                                 int my, int mz) {
  GridRay g;
  g.worldOrg = worldOrg;
  g.worldDir = worldDir;
  const V3 org = worldOrg - blo;
  const V3 ext = bhi - blo;
  g.worldToUnit = v3(1.f / ext.x, 1.f / ext.y, 1.f / ext.z);
  const V3 unitToGrid = v3(float(mx), float(my), float(mz));
  g.org = unitToGrid * g.worldToUnit * org;
  const V3 dir = g.worldToUnit * unitToGrid * worldDir;
  g.worldToGridT = sqrtf(dot3(dir, dir));
  g.gridToWorldT = 1.f / g.worldToGridT;
  g.dir = normalize3(dir);
  return g;
}

// Root-box clip (deviceCode.cu:1676-1726) with ray tmin 0 and tmax 1e20.
__device__ __forceinline__ bool clip_root(V3 org, V3 dir, V3 lb, V3 rt, float& t0, float& t1) {
  const float dfx = 1.0f / dir.x, dfy = 1.0f / dir.y, dfz = 1.0f / dir.z;
  const float a1 = (lb.x - org.x) * dfx, a2 = (rt.x - org.x) * dfx;
  const float a3 = (lb.y - org.y) * dfy, a4 = (rt.y - org.y) * dfy;
  const float a5 = (lb.z - org.z) * dfz, a6 = (rt.z - org.z) * dfz;
  float thit0 = fmaxf(fmaxf(fminf(a1, a2), fminf(a3, a4)), fminf(a5, a6));
  const float thit1 = fminf(fminf(fmaxf(a1, a2), fmaxf(a3, a4)), fmaxf(a5, a6));
  if (thit1 < 0) return false;
  if (thit0 >= thit1) return false;
  thit0 = fmaxf(thit0, 0.f);
  t0 = fmaxf(0.f, thit0);
  t1 = fminf(1e20f, thit1);
  return true;
}

// Camera ray (deviceCode.cu:293-300 jitter, generateRay 61-76).
__device__ __forceinline__ void camera_ray(const CameraParams& cam, int px, int py, int W, int H,
                                           float jx, float jy, V3& org, V3& dir) {
  const float su = (float(px) + jx) / float(W);
  const float sv = (float(py) + jy) / float(H);
  org = load3(cam.org);
  V3 d = load3(cam.llc) + su * load3(cam.hor) + sv * load3(cam.ver);
  if (fabsf(d.x) < 1e-5f) d.x = 1e-5f;
  if (fabsf(d.y) < 1e-5f) d.y = 1e-5f;
  if (fabsf(d.z) < 1e-5f) d.z = 1e-5f;
  dir = normalize3(d - org);
}

}  // namespace mdwt
