// CPU oracle for multi-density Woodcock tracking. TEST INFRASTRUCTURE ONLY:
// loaded by tests/, __graft_entry__.smoke() and bench.py's CPU baseline leg as
// the checker; the HIP product never links or calls it.
//
// Scalar restatement of the reference's RAW render path. Each function cites
// the reference file:line it follows. Build: see oracle/Makefile
// (g++ -O2 -ffp-contract=off, no fast-math, so every fp32 operation is rounded
// exactly once in source order, like the HIP kernels built with
// -ffp-contract=off).
#include "mdwt_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kMaxChannels = 12;  // deviceCode.h:13 MAX_CHANNELS

// ---------------------------------------------------------------- vectors ---
// OWL vec3f/vec4f stand-ins. Binary vector ops are component-wise and
// evaluated left to right, as C++ evaluates the reference's expressions.
struct vec3f { float x, y, z; };
struct vec4f { float x, y, z, w; };
struct vec3i { int x, y, z; };

inline vec3f operator+(vec3f a, vec3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3f operator-(vec3f a, vec3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3f operator*(vec3f a, vec3f b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3f operator*(float s, vec3f a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3f operator*(vec3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3f operator/(float s, vec3f a) { return {s / a.x, s / a.y, s / a.z}; }
inline vec3f operator/(vec3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline vec3f operator-(vec3f a) { return {-a.x, -a.y, -a.z}; }
inline float dot(vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(vec3f a) { return sqrtf(dot(a, a)); }
inline vec3f normalize(vec3f a) { return a * (1.f / sqrtf(dot(a, a))); }
inline vec3f cross(vec3f a, vec3f b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline vec3f vmin(vec3f a, vec3f b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
inline vec3f vmax(vec3f a, vec3f b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
inline vec3f vabs(vec3f a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
inline vec3f vfloor(vec3f a) { return {floorf(a.x), floorf(a.y), floorf(a.z)}; }
inline float reduce_min(vec3f a) { return fminf(fminf(a.x, a.y), a.z); }
inline float reduce_max(vec3f a) { return fmaxf(fmaxf(a.x, a.y), a.z); }

inline vec4f operator+(vec4f a, vec4f b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline vec4f operator*(float s, vec4f a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
inline vec4f operator/(vec4f a, float s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }

// ------------------------------------------------------------------- RNG ----
// Philox4x32 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
// as 1, 2, 3", SC'11) with 7 rounds, the fewest for which that paper reports
// Philox4x32 passing BigCrush (the round function and key schedule are those
// of the ten-round generator, whose published known answers pin them).
// Replaces OWL's LCG<4> (deviceCode.h:145, deviceCode.cu:296-297), whose
// streams are keyed by threadIdx; here a stream is keyed by (seed, pixel) and
// counted by (frame, sample, block, stream).
constexpr int kPhiloxRounds = 7;
inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
  const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
  const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
  const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}
inline void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4],
                   int rounds = kPhiloxRounds) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < rounds; ++r) {
    if (r > 0) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    philox_round(c, k);
  }
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// One stream of uniforms in [0,1) per (pixel, frame, sample, stream): Philox
// blocks under key (seed, pixel), words converted as (w >> 8) * 2^-24. Fixed
// word layout: the pixel jitter is word 0 of block (frame, sample, 0, stream)
// (and with two-axis jitter, oracle_set_jitter_2d, the y jitter word 1).
//
// -m 0 and -m 2 track each (macrocell visit v, channel c) from the cell's
// entry (v counts the cells of the ray's DDA walk, the same cells for every
// channel pass of -m 0); -m 3 / 4 / 5 track their one summed majorant per
// visit v with the positions of channel 0:
//   * the first free-flight distance of every channel of group g = c / 4 at
//     visit v is word c mod 4 of the cell block (frame, sample,
//     1 + 16 v + 12 + g, stream): one block serves up to four channels, and
//     most (v, c) steps leave the cell at that first step;
//   * after a step j lands inside the cell, pair j of the (v, c) stream holds
//     its null-collision threshold and the next step's distance: words
//     2 (j mod 2) and 2 (j mod 2) + 1 of block (frame, sample, 1 + 16 v + c,
//     stream | 2 floor(j / 2)).
// -m 1 runs one exponential clock per channel in a cell on the (v, c)
// positions of -m 2: the clock of channel c starts at its first distance and
// its j-th event uses pair j of the (v, c) stream. -m 4 / 5 track their summed
// majorant per visit like -m 3 (channel 0's positions).
// step() / thr() count free-flight steps from the ray's start (step j: words
// 2 (j mod 2) and 2 (j mod 2) + 1 of block (frame, sample, 1, stream | 2
// floor(j / 2))); no mode uses that layout now, and in the faithful path they
// take the running stream.
// The reference draws the same quantities from one sequential LCG stream
// (deviceCode.cu:296-298, 563, 583, 797, 830, 913, 931); fixed positions let
// the GPU compute one Philox block for all first steps of a cell, and draw
// any (v, c) step independently of the others.
//
// Sequential mode (the reference-faithful path, oracle_set_faithful): seek()
// does nothing and every draw after the jitter takes the next word of one
// running stream (word i & 3 of block (frame, sample, 1 + i / 4, stream)),
// in the order the reference calls prd.rng(): each step's distance draw, and
// its threshold draw only when the step lands inside the cell.
struct Rng {
  uint32_t key[2];
  uint32_t frame, sample, stream;
  uint32_t buf[4];
  uint32_t c2 = 0, c3 = 0;  // counter words of the block held in buf
  bool held = false;
  bool seq = false;   // one running stream (reference draw order)
  uint32_t n = 0;     // next word of the running stream
  uint32_t k = 0;     // next free-flight step since the last seek
  uint32_t cur = 0;   // current step
  uint32_t base = 0;  // block offset set by seek
  Rng(uint32_t seed, uint32_t pixel, uint32_t frame_, uint32_t sample_, uint32_t stream_)
      : frame(frame_), sample(sample_), stream(stream_) {
    key[0] = seed; key[1] = pixel;
  }
  const uint32_t* block(uint32_t w2, uint32_t w3) {
    if (!held || w2 != c2 || w3 != c3) {
      const uint32_t ctr[4] = {frame, sample, w2, w3};
      philox(ctr, key, buf);
      c2 = w2; c3 = w3; held = true;
    }
    return buf;
  }
  static float unit(uint32_t w) { return float(w >> 8) * 0x1p-24f; }  // exact
  // deviceCode.cu:297-298 adds one uniform to both pixel coordinates; with
  // two_d the y jitter is the block's next word
  void jitter2(bool two_d, float& jx, float& jy) {
    const uint32_t* w = block(0u, stream);
    jx = unit(w[0]);
    jy = two_d ? unit(w[1]) : jx;
  }
  void seek(uint32_t b) {
    if (seq) return;
    base = b;
    k = 0;
  }
  float next() {
    const uint32_t i = n++;
    return unit(block(1u + (i >> 2), stream)[i & 3u]);
  }
  // -m 0 / -m 2: first distance of channel c at visit v; starts the (v, c)
  // stream's pairs
  float first_dist(uint32_t v, uint32_t c) {
    if (seq) return next();
    base = (v << 4) | c;
    cur = 0;
    return unit(block(1u + ((v << 4) | (12u + (c >> 2))), stream)[c & 3u]);
  }
  // threshold of the current step (pair cur, word 0)
  float pair_thr() {
    if (seq) return next();
    return unit(block(1u + base, stream | ((cur >> 1) << 1))[2 * (cur & 1u)]);
  }
  // distance of the next step (pair cur, word 1); advances to that step
  float pair_dist() {
    if (seq) return next();
    const uint32_t p = cur++;
    return unit(block(1u + base, stream | ((p >> 1) << 1))[2 * (p & 1u) + 1]);
  }
  // distance uniform of the next free-flight step
  float step() {
    if (seq) return next();
    cur = k++;
    return unit(block(1u + base, stream | ((cur >> 1) << 1))[2 * (cur & 1u)]);
  }
  // threshold uniform of the current step
  float thr() {
    if (seq) return next();
    return unit(block(1u + base, stream | ((cur >> 1) << 1))[2 * (cur & 1u) + 1]);
  }
  // -m 1: the first distance of channel c at visit v (no pair state) and
  // pair j of the (v, c) stream (its threshold and the next distance)
  float cell_word(uint32_t v, uint32_t c) {
    return unit(block(1u + ((v << 4) | (12u + (c >> 2))), stream)[c & 3u]);
  }
  void vc_pair(uint32_t v, uint32_t c, uint32_t j, float& th, float& dist) {
    const uint32_t* b = block(1u + ((v << 4) | c), stream | ((j >> 1) << 1));
    th = unit(b[2 * (j & 1u)]);
    dist = unit(b[2 * (j & 1u) + 1]);
  }
};

// --------------------------------------------------------------- logf -------
// Portable fp32 natural log for x > 0 normal: x = 2^e m, m in [sqrt(1/2),
// sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1), atanh series to s^9
// (Abramowitz & Stegun 4.1.27); e ln2 split in a Cody-Waite hi/lo pair.
// Used in place of log() at deviceCode.cu:563,797,913 so that CPU and GPU
// free-flight distances round identically.
inline float plogf(float x) {
  uint32_t bits;
  std::memcpy(&bits, &x, 4);
  int e = int((bits >> 23) & 0xffu) - 127;
  uint32_t mb = (bits & 0x007fffffu) | 0x3f800000u;
  float m;
  std::memcpy(&m, &mb, 4);
  if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  const float p = ((((0.11111111f * z + 0.14285715f) * z + 0.2f) * z + 0.33333334f) * z);
  const float r = (s + s) + (s + s) * p;
  const float fe = float(e);
  return fe * 0.693145751953125f + (r + fe * 1.42860677e-06f);
}

// The same reduction with log m = f P(f), f = m - 1, P a degree-9 polynomial
// (least-squares fit of log1p(f)/f on Chebyshev nodes of [sqrt(1/2)-1,
// sqrt(2)-1], P(0) = 1) in Horner form with fused multiply-adds: no division,
// so fewer instructions per free-flight step. fmaf rounds once on the host and
// on the GPU, so both sides agree bit for bit. Used by every free-flight step
// (deviceCode.cu:913, 805, 676, 739, 564, 1023); plogf remains for pow().
inline float plogf_poly(float x) {
  uint32_t bits;
  std::memcpy(&bits, &x, 4);
  int e = int((bits >> 23) & 0xffu) - 127;
  uint32_t mb = (bits & 0x007fffffu) | 0x3f800000u;
  float m;
  std::memcpy(&m, &mb, 4);
  if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
  const float f = m - 1.0f;
  float p = -7.637512684e-02f;
  p = std::fmaf(p, f, 1.291354448e-01f);
  p = std::fmaf(p, f, -1.324101686e-01f);
  p = std::fmaf(p, f, 1.418020874e-01f);
  p = std::fmaf(p, f, -1.660911888e-01f);
  p = std::fmaf(p, f, 2.000206262e-01f);
  p = std::fmaf(p, f, -2.500158548e-01f);
  p = std::fmaf(p, f, 3.333332539e-01f);
  p = std::fmaf(p, f, -4.999998808e-01f);
  p = std::fmaf(p, f, 1.000000000e+00f);
  const float r = f * p;
  const float fe = float(e);
  return fe * 0.693145751953125f + (r + fe * 1.42860677e-06f);
}

// Portable fp32 exp: x = n ln2 + r (Cody-Waite, same ln2 split as plogf),
// e^r by its Taylor polynomial to r^6 (|r| <= ln2/2, truncation < 2^-24
// relative), scaled by 2^n built from the exponent bits. Flushes to 0 below
// -87.3 (no denormal results) and saturates to +inf above 88 (so that 2^n
// stays a normal number).
inline float pexpf(float x) {
  if (!(x > -87.3f)) return 0.0f;
  if (x > 88.0f) return INFINITY;
  const float n = floorf(x * 1.44269504f + 0.5f);
  const float r = (x - n * 0.693145751953125f) - n * 1.42860677e-06f;
  const float p =
      1.0f + r * (1.0f + r * (0.5f + r * (0.16666667f + r * (0.041666668f + r * (0.008333334f +
                                                                           r * 0.0013888889f)))));
  const int ni = int(n);
  const uint32_t b = uint32_t(ni + 127) << 23;
  float scale;
  std::memcpy(&scale, &b, 4);
  return p * scale;
}
// pow(x, y) for x > 0 as exp(y log x) with the portable pair above; replaces
// pow() at deviceCode.cu:383 so that CPU and GPU round identically.
inline float ppowf(float x, float y) { return pexpf(y * plogf(x)); }

// ------------------------------------------------------------- textures ----
// Clamp-to-edge linear lookup of a 1D float4 table at normalized u: texel
// coordinate x = u*N - 0.5, lerp a + f*(b-a) (renderer.cpp:1140-1152 texture
// descriptor: linear filter, clamp, normalized coordinates).
inline int clamp_index(float fl, int n) {
  // clamp in float first so that int conversion is always defined
  float c = fminf(fmaxf(fl, -1.0f), float(n));
  int i = int(c);
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}
inline int clamp_index_plus1(float fl, int n) {
  float c = fminf(fmaxf(fl, -1.0f), float(n));
  int i = int(c) + 1;
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

struct TF {
  std::vector<vec4f> texels;  // alpha already cubed (renderer.cpp:1098-1101)
  float opacity_scale = 1.f;
  float xf_lo = 0.f, xf_hi = 1.f;
  bool set = false;
};

inline vec4f tf_lookup_u(const TF& tf, float u) {
  const int n = int(tf.texels.size());
  const float x = u * float(n) - 0.5f;
  const float fl = floorf(x);
  const float f = x - fl;
  const vec4f a = tf.texels[clamp_index(fl, n)];
  const vec4f b = tf.texels[clamp_index_plus1(fl, n)];
  return {a.x + f * (b.x - a.x), a.y + f * (b.y - a.y), a.z + f * (b.z - a.z),
          a.w + f * (b.w - a.w)};
}

inline float half_to_float(uint16_t h) {
  const uint32_t s = uint32_t(h >> 15) << 31;
  uint32_t e = (h >> 10) & 0x1f;
  uint32_t m = h & 0x3ff;
  uint32_t bits;
  if (e == 0) {
    if (m == 0) {
      bits = s;
    } else {  // subnormal half -> normal float
      int ee = -1;
      do { m <<= 1; ++ee; } while ((m & 0x400) == 0);
      m &= 0x3ff;
      bits = s | (uint32_t(127 - 15 - ee) << 23) | (m << 13);
    }
  } else if (e == 31) {
    bits = s | 0x7f800000u | (m << 13);
  } else {
    bits = s | ((e + 127 - 15) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

struct Channel {
  int dtype = 0;
  int dims[3] = {0, 0, 0};
  std::vector<uint8_t> raw;
  float scale = 1.f, offset = 0.f;  // value = raw * scale + offset (rawFile.cpp:100-123)
  float vmin = 0.f, vmax = 0.f;     // converted min/max
  float raw_at(size_t i) const {
    switch (dtype) {
      case ORACLE_U8: return float(raw[i]);
      case ORACLE_U16: { uint16_t v; std::memcpy(&v, &raw[2 * i], 2); return float(v); }
      case ORACLE_F16: { uint16_t v; std::memcpy(&v, &raw[2 * i], 2); return half_to_float(v); }
      default: { float v; std::memcpy(&v, &raw[4 * i], 4); return v; }
    }
  }
  // converted voxel value (rawFile.cpp:100-123): u8 -> v*255, u16 -> v*65535
  float value_at(size_t i) const { return raw_at(i) * scale + offset; }
};

// renderer.h:134-141 MaterialData defaults
struct Material {
  float diffuse = 0.8f, specular = 0.3f, shininess = 8.0f, gradient_step = 0.01f;
  vec3f spec{1.f, 1.f, 1.f};
};

struct Hit {
  bool missed = true;
  vec4f rgba{0, 0, 0, 0};
  float tHit = 0.f;
  int channel = 0;
  uint64_t samples = 0, rejections = 0, visits = 0;
};

}  // namespace

struct oracle_renderer {
  std::string err;
  std::vector<Channel> ch;
  vec3f remap{1.f, 1.f, 1.f};
  bool remap_set = false;
  bool inited = false;
  int mode = 0;
  int mc[3] = {0, 0, 0};
  vec3f blo{0, 0, 0}, bhi{1, 1, 1};       // world box (rawFile.h:162, reshapeBounds)
  std::vector<float> ranges;              // float2 per (cell, ch)
  std::vector<float> maj;
  std::vector<TF> tf;
  std::vector<float> dom_lo, dom_hi;      // value domains (renderer.cpp:743-745)
  // camera (camera.h state) and derived launch values (renderer.cpp:1058-1092)
  vec3f cam_pos{0, -1, 0};
  vec3f fvx{1, 0, 0}, fvy{0, 1, 0}, fvz{0, 0, 1};
  vec3f cam_up{0, 1, 0};
  float fovy = 60.f;
  vec3f c_org{0, 0, 0}, c_llc{0, 0, 0}, c_hor{0, 0, 0}, c_ver{0, 0, 0};
  float dt = 0.05f;
  int W = 1024, H = 1024, spp = 1;
  bool accumulate = true;
  vec3f bg{0, 0, 0};
  float light_intensity = 1.0f, ambient = 0.1f;   // renderer.h:124-126
  vec3f light_dir{0.0001f, -1.f, 0.0001f};         // renderer.h:124
  bool enable_shadows = false, enable_gradient = false;  // renderer.h:121-122
  Material mat[kMaxChannels];                      // renderer.h:133-142
  uint32_t seed = 0x5eed1234u;
  uint32_t frame_id = 0, accum_id = 0;
  std::vector<vec4f> accum, m2;
  uint64_t counters[4] = {0, 0, 0, 0};
  int build_threads = 1;
  // reference-faithful tracking (oracle_set_faithful): libm log, the two
  // divisions of the TF remap, one sequential RNG stream per ray
  bool faithful = false;
  bool jitter2d = false;  // oracle_set_jitter_2d
  bool conservative_maj = false;  // oracle_set_conservative_majorants
};

namespace {

int fail(oracle_renderer* o, const std::string& m) {
  o->err = m;
  return 1;
}

// ------------------------------------------------------------ camera -------
// camera.h:83-110 setOrientation (+ forceUpFrame camera.h:72-80)
void cam_set_orientation(oracle_renderer* o, vec3f origin, vec3f interest, vec3f up,
                         float fovy) {
  o->fovy = fovy;
  o->cam_pos = origin;
  o->cam_up = up;
  const bool same = origin.x == interest.x && origin.y == interest.y && origin.z == interest.z;
  o->fvz = same ? vec3f{0, 0, 1} : -normalize(interest - origin);
  o->fvx = cross(up, o->fvz);
  if (dot(o->fvx, o->fvx) < 1e-8f)
    o->fvx = vec3f{0, 1, 0};
  else
    o->fvx = normalize(o->fvx);
  o->fvy = normalize(cross(o->fvz, o->fvx));
  // forceUpFrame
  if (fabsf(dot(o->fvz, o->cam_up)) < 1e-6f) return;
  o->fvx = normalize(cross(o->cam_up, o->fvz));
  o->fvy = normalize(cross(o->fvz, o->fvx));
}

inline float to_radian(float d) { return d * float(M_PI / 180.0); }
inline float to_degrees(float r) { return r * float(180.0 / M_PI); }

// renderer.cpp:1058-1092 UpdateCamera
void update_camera(oracle_renderer* o) {
  const vec3f lookFrom = o->cam_pos;
  const vec3f lookAt = o->cam_pos - o->fvz;  // camera.h:59 getAt
  const vec3f lookUp = o->fvy;               // camera.h:61 getUp
  const float cosFovy = cosf(to_radian(o->fovy));
  const float vfov = to_degrees(acosf(cosFovy));
  const vec3f vup = lookUp;
  const float aspect = float(o->W) / float(o->H);
  const float theta = vfov * float(M_PI) / 180.0f;
  const float half_height = tanf(theta / 2.0f);
  const float half_width = aspect * half_height;
  const float focusDist = 10.f;
  const vec3f origin = lookFrom;
  const vec3f w = normalize(lookFrom - lookAt);
  const vec3f u = normalize(cross(vup, w));
  const vec3f v = cross(w, u);
  o->c_org = origin;
  o->c_llc = origin - half_width * focusDist * u - half_height * focusDist * v - focusDist * w;
  o->c_hor = 2.0f * half_width * focusDist * u;
  o->c_ver = 2.0f * half_height * focusDist * v;
  o->accum_id = 0;  // ResetAccumulation (renderer.cpp:1091)
}

// ------------------------------------------------------ range grid ---------
// renderer.cu:387-391 project
inline int project(float f, float lo, float hi, int dim) {
  const float q = float(dim) * (f - lo) / (hi - lo);
  const int i = int(q);
  return std::max(0, std::min(dim - 1, i));
}

// renderer.cu:1116-1179 buildSpatialMacrocells (RAW), sizeMCs 330-347,
// rasterElements (RAW) 835-893, rasterBox (float2) 487-519.
void build_ranges(oracle_renderer* o) {
  const int C = int(o->ch.size());
  const int Mx = o->mc[0], My = o->mc[1], Mz = o->mc[2];
  const size_t ncell = size_t(Mx) * My * Mz;
  o->ranges.assign(ncell * C * 2, 0.f);
  for (size_t i = 0; i < ncell * C; ++i) {  // sizeMCs: (1e20, -1e20)
    o->ranges[2 * i] = 1e20f;
    o->ranges[2 * i + 1] = -1e20f;
  }
  const vec3f wlo = o->blo, whi = o->bhi;
  for (int c = 0; c < C; ++c) {
    const Channel& chn = o->ch[c];
    const int nx = chn.dims[0], ny = chn.dims[1], nz = chn.dims[2];
    const size_t nvox = size_t(nx) * ny * nz;
    std::vector<float> val(nvox);  // the reference's float scalar buffer (renderer.cpp:690-696)
    for (size_t i = 0; i < nvox; ++i) val[i] = chn.value_at(i);
    const vec3f size = whi - wlo;
    const vec3f boxLen = {size.x / float(nx), size.y / float(ny), size.z / float(nz)};
    // One thread per voxel in the reference, merged by order-independent CAS
    // min/max; here z-slabs per OpenMP thread, each into a private grid that
    // is min/max-merged afterwards (same result for any thread count).
    const int nthr = o->build_threads > 0 ? o->build_threads : 1;
    std::vector<std::vector<float>> part(nthr);
#ifdef _OPENMP
#pragma omp parallel num_threads(nthr)
#endif
    {
#ifdef _OPENMP
      const int tid = omp_get_thread_num();
#else
      const int tid = 0;
#endif
      std::vector<float>& g = part[tid];
      g.assign(ncell * 2, 0.f);
      for (size_t i = 0; i < ncell; ++i) {
        g[2 * i] = 1e20f;
        g[2 * i + 1] = -1e20f;
      }
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
      for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
          for (int ix = 0; ix < nx; ++ix) {
            const size_t idx = size_t(ix) + size_t(iy) * nx + size_t(iz) * nx * ny;
            const vec3f vxlLower = wlo + vec3f{float(ix), float(iy), float(iz)} * boxLen;
            const float s = val[idx];
            vec3f lower = vxlLower, upper = vxlLower + boxLen;
            float lw = s, uw = s;
            for (int dz = -1; dz <= 1; ++dz)
              for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                  const int qx = ix + dx, qy = iy + dy, qz = iz + dz;
                  if (qx < 0 || qx >= nx) continue;
                  if (qy < 0 || qy >= ny) continue;
                  if (qz < 0 || qz >= nz) continue;
                  const size_t q = size_t(qx) + size_t(qy) * nx + size_t(qz) * nx * ny;
                  const vec3f nl = wlo + vec3f{float(qx), float(qy), float(qz)} * boxLen;
                  const float sn = val[q];
                  lower = vmin(lower, nl);
                  upper = vmax(upper, nl);
                  lw = fminf(lw, sn);
                  uw = fmaxf(uw, sn);
                }
            // rasterBox (float2)
            if (lower.x >= upper.x || lower.y >= upper.y || lower.z >= upper.z) continue;
            const int lx = project(lower.x, wlo.x, whi.x, Mx);
            const int ly = project(lower.y, wlo.y, whi.y, My);
            const int lz = project(lower.z, wlo.z, whi.z, Mz);
            const int hx = std::min(project(upper.x, wlo.x, whi.x, Mx), Mx - 1);
            const int hy = std::min(project(upper.y, wlo.y, whi.y, My), My - 1);
            const int hz = std::min(project(upper.z, wlo.z, whi.z, Mz), Mz - 1);
            for (int cz = lz; cz <= hz; ++cz)
              for (int cy = ly; cy <= hy; ++cy)
                for (int cx = lx; cx <= hx; ++cx) {
                  const size_t cell = size_t(cx) + size_t(cy) * Mx + size_t(cz) * Mx * My;
                  float* r = &g[2 * cell];
                  if (lw < r[0]) r[0] = lw;  // CAS atomicMin (renderer.cu:362-371)
                  if (uw > r[1]) r[1] = uw;  // CAS atomicMax (renderer.cu:374-384)
                }
          }
    }
    for (int t = 0; t < nthr; ++t)
      for (size_t cell = 0; cell < ncell; ++cell) {
        float* r = &o->ranges[2 * (cell * C + c)];
        const float* p = &part[t][2 * cell];
        if (p[0] < r[0]) r[0] = p[0];
        if (p[1] > r[1]) r[1] = p[1];
      }
  }
}

// renderer.cu:159-203 (_recalculateDensityRangesMM) per-channel majorant of
// one (cell, channel); renderer.cu:111-157 sums the same quantity for mode 3.
float channel_majorant(const oracle_renderer* o, int c, float mn, float mx) {
  const TF& tf = o->tf[c];
  const float vlo = o->dom_lo[c], vhi = o->dom_hi[c];
  const int n = int(tf.texels.size());
  const float remappedMin1 = (mn - vlo) / (vhi - vlo);
  const float remappedMin = (remappedMin1 - tf.xf_lo) / (tf.xf_hi - tf.xf_lo);
  const float remappedMax1 = (mx - vlo) / (vhi - vlo);
  const float remappedMax = (remappedMax1 - tf.xf_lo) / (tf.xf_hi - tf.xf_lo);
  const float addr1 = remappedMin * float(n);
  const float addr2 = remappedMax * float(n);
  const int addrMin = clamp_index(fminf(floorf(addr1), floorf(addr2)), n);
  const int addrMax = clamp_index(fmaxf(ceilf(addr1), ceilf(addr2)), n);
  float m = 0.f;
  for (int i = addrMin; i <= addrMax; ++i) {
    const float d = tf_lookup_u(tf, (float(i) + 0.5f) / float(n)).w * tf.opacity_scale;
    m = (i == addrMin) ? d : fmaxf(m, d);
  }
  float d = tf_lookup_u(tf, (addr1 + 0.5f) / float(n)).w * tf.opacity_scale;
  m = fmaxf(m, d);
  d = tf_lookup_u(tf, (addr2 + 0.5f) / float(n)).w * tf.opacity_scale;
  m = fmaxf(m, d);
  return m;
}

// The conservative option (oracle_set_conservative_majorants), no reference
// counterpart: the largest opacity of every texel a lookup of a value in
// [mn, mx] reads, with the lookup interval widened by 2^-12 (1 + |r|) on each
// side and the maximum scaled by 1 + 2^-20. A lookup at u reads texels
// floor(u n - 0.5) and the next one (tf_lookup_u).
float channel_majorant_conservative(const oracle_renderer* o, int c, float mn, float mx) {
  const TF& tf = o->tf[c];
  const float vlo = o->dom_lo[c], vhi = o->dom_hi[c];
  const int n = int(tf.texels.size());
  const float r1 = ((mn - vlo) / (vhi - vlo) - tf.xf_lo) / (tf.xf_hi - tf.xf_lo);
  const float r2 = ((mx - vlo) / (vhi - vlo) - tf.xf_lo) / (tf.xf_hi - tf.xf_lo);
  const float a = fminf(r1, r2), b = fmaxf(r1, r2);
  const float lo = a - 0x1p-12f * (1.0f + fabsf(a)), hi = b + 0x1p-12f * (1.0f + fabsf(b));
  const int i0 = clamp_index(floorf(lo * float(n) - 0.5f), n);
  const int i1 = clamp_index_plus1(floorf(hi * float(n) - 0.5f), n);
  float m = 0.f;
  for (int i = i0; i <= i1; ++i) m = fmaxf(m, tf.texels[i].w);
  return m * tf.opacity_scale * (1.0f + 0x1p-20f);
}

// renderer.cu:206-274 RecalculateDensityRanges (grid part)
void build_majorants(oracle_renderer* o) {
  const int C = int(o->ch.size());
  const size_t ncell = size_t(o->mc[0]) * o->mc[1] * o->mc[2];
  const bool per_channel = o->mode <= 2;
  o->maj.assign(per_channel ? ncell * C : ncell, 0.f);
  for (size_t cell = 0; cell < ncell; ++cell) {
    float sum = 0.f;
    for (int c = 0; c < C; ++c) {
      const float mn = o->ranges[2 * (cell * C + c)];
      const float mx = o->ranges[2 * (cell * C + c) + 1];
      if (mx < mn) continue;  // empty box
      const float m = o->conservative_maj ? channel_majorant_conservative(o, c, mn, mx)
                                          : channel_majorant(o, c, mn, mx);
      if (per_channel)
        o->maj[cell * C + c] = m;
      else
        sum += m;
    }
    if (!per_channel) o->maj[cell] = sum;
  }
}

bool tfs_ready(const oracle_renderer* o) {
  for (const TF& t : o->tf)
    if (!t.set) return false;
  return true;
}

// ------------------------------------------------------ device programs ----
// deviceCode.cu:27-42 transferFunction. Its two divisions by per-channel
// constants ((f - lo) / (hi - lo), then by the TF range width) are products
// with the correctly rounded reciprocals of those constants here, as on the
// GPU: the remapped value moves by about an ulp, far below one texel.
// The reference-faithful path keeps the two divisions as written.
inline vec4f transfer_function(const oracle_renderer* o, float f, int c) {
  const float vlo = o->dom_lo[c], vhi = o->dom_hi[c];
  if (f < vlo || f > vhi) return {1.f, 0.f, 1.f, 0.f};
  const TF& tf = o->tf[c];
  float r2;
  if (o->faithful) {
    r2 = ((f - vlo) / (vhi - vlo) - tf.xf_lo) / (tf.xf_hi - tf.xf_lo);
  } else {
    const float idw = 1.0f / (vhi - vlo), ixw = 1.0f / (tf.xf_hi - tf.xf_lo);
    const float r1 = (f - vlo) * idw;
    r2 = (r1 - tf.xf_lo) * ixw;
  }
  vec4f xf = tf_lookup_u(tf, r2);
  xf.w = xf.w * tf.opacity_scale;
  return xf;
}

// log(1 - u) of a free-flight step (deviceCode.cu:563, 805, 913): the
// division-free polynomial shared with the GPU, or libm's logf on the
// reference-faithful path.
inline float flight_log(const oracle_renderer* o, float x) {
  return o->faithful ? std::log(x) : plogf_poly(x);
}

// deviceCode.cu:122-132 sampleVolumeTexture -> sample 91-120 (one texture):
// NaN outside [0,1]^3, else clamp-to-edge trilinear at texel coordinate p*N-0.5.
inline float sample_volume(const oracle_renderer* o, vec3f p, int c) {
  if (p.x < 0.0f || p.x > 1.0f || p.y < 0.0f || p.y > 1.0f || p.z < 0.0f || p.z > 1.0f)
    return NAN;
  const Channel& chn = o->ch[c];
  const int nx = chn.dims[0], ny = chn.dims[1], nz = chn.dims[2];
  const float x = p.x * float(nx) - 0.5f;
  const float y = p.y * float(ny) - 0.5f;
  const float z = p.z * float(nz) - 0.5f;
  const float flx = floorf(x), fly = floorf(y), flz = floorf(z);
  const float fx = x - flx, fy = y - fly, fz = z - flz;
  const int x0 = clamp_index(flx, nx), x1 = clamp_index_plus1(flx, nx);
  const int y0 = clamp_index(fly, ny), y1 = clamp_index_plus1(fly, ny);
  const int z0 = clamp_index(flz, nz), z1 = clamp_index_plus1(flz, nz);
  auto at = [&](int i, int j, int k) {
    return chn.raw_at(size_t(i) + size_t(j) * nx + size_t(k) * size_t(nx) * ny);
  };
  const float v000 = at(x0, y0, z0), v100 = at(x1, y0, z0);
  const float v010 = at(x0, y1, z0), v110 = at(x1, y1, z0);
  const float v001 = at(x0, y0, z1), v101 = at(x1, y0, z1);
  const float v011 = at(x0, y1, z1), v111 = at(x1, y1, z1);
  const float c00 = v000 + fx * (v100 - v000);
  const float c10 = v010 + fx * (v110 - v010);
  const float c01 = v001 + fx * (v101 - v001);
  const float c11 = v011 + fx * (v111 - v011);
  const float c0 = c00 + fy * (c10 - c00);
  const float c1 = c01 + fy * (c11 - c01);
  const float r = c0 + fz * (c1 - c0);
  return r * chn.scale + chn.offset;
}

// Grid-space ray set-up shared by the closest-hit programs
// (deviceCode.cu:866-888; identical at 521-543 and 755-777).
struct GridRay {
  vec3f worldOrg, worldDir, worldToUnit, org, dir;
  float worldToGridT, gridToWorldT;
};
inline GridRay grid_ray(const oracle_renderer* o, vec3f worldOrg, vec3f worldDir) {
  GridRay g;
  g.worldOrg = worldOrg;
  g.worldDir = worldDir;
  vec3f org = worldOrg - o->blo;
  g.worldToUnit = 1.f / (o->bhi - o->blo);
  const vec3f unitToGrid = {float(o->mc[0]), float(o->mc[1]), float(o->mc[2])};
  g.org = unitToGrid * g.worldToUnit * org;
  vec3f dir = g.worldToUnit * unitToGrid * worldDir;
  g.worldToGridT = length(dir);
  g.gridToWorldT = 1.f / g.worldToGridT;
  g.dir = normalize(dir);
  return g;
}

// dda.h:82-203 dda3. The lambda returns true to continue. Reproduces the
// reference's zero-direction handling (dda.h:96-116: t_nr written twice,
// t_fr untouched), which drops rays with an exactly zero grid direction.
template <typename F>
void dda3(vec3f org, vec3f dir, float tMax, const int grid[3], F&& lambda) {
  const vec3f f_size = {float(grid[0]), float(grid[1]), float(grid[2])};
  const vec3f rcp_dir = 1.f / dir;
  const vec3f t_lo = (vec3f{0.f, 0.f, 0.f} - org) * rcp_dir;
  const vec3f t_hi = (f_size - org) * rcp_dir;
  vec3f t_nr = vmin(t_lo, t_hi);
  const vec3f t_fr = vmax(t_lo, t_hi);
  if (dir.x == 0.f) {
    if (org.x < 0.f || org.x > f_size.x) return;
    t_nr.x = INFINITY;
  }
  if (dir.y == 0.f) {
    if (org.y < 0.f || org.y > f_size.y) return;
    t_nr.y = INFINITY;
  }
  if (dir.z == 0.f) {
    if (org.z < 0.f || org.z > f_size.z) return;
    t_nr.z = INFINITY;
  }
  const float ray_t0 = fmaxf(0.f, reduce_max(t_nr));
  const float ray_t1 = fminf(tMax, reduce_min(t_fr));
  if (ray_t0 > ray_t1) return;
  const vec3f org_in_volume = org + ray_t0 * dir;
  const vec3f f_cell = vmax(vec3f{0.f, 0.f, 0.f},
                            vmin(f_size - vec3f{1.f, 1.f, 1.f}, vfloor(org_in_volume)));
  const vec3f f_cell_end = {dir.x > 0.f ? f_cell.x + 1.f : f_cell.x,
                            dir.y > 0.f ? f_cell.y + 1.f : f_cell.y,
                            dir.z > 0.f ? f_cell.z + 1.f : f_cell.z};
  const vec3f t_step = vabs(rcp_dir);
  vec3f t_next = {
      (dir.x == 0.f) ? INFINITY : fabsf(f_cell_end.x - org_in_volume.x) * t_step.x,
      (dir.y == 0.f) ? INFINITY : fabsf(f_cell_end.y - org_in_volume.y) * t_step.y,
      (dir.z == 0.f) ? INFINITY : fabsf(f_cell_end.z - org_in_volume.z) * t_step.z};
  const vec3i stop = {dir.x > 0.f ? grid[0] : -1, dir.y > 0.f ? grid[1] : -1,
                      dir.z > 0.f ? grid[2] : -1};
  const vec3i delta = {dir.x > 0.f ? 1 : -1, dir.y > 0.f ? 1 : -1, dir.z > 0.f ? 1 : -1};
  vec3i cell = {int(f_cell.x), int(f_cell.y), int(f_cell.z)};
  float next_cell_begin = 0.f;
  while (true) {
    const float t_closest = reduce_min(t_next);
    const float cell_t0 = ray_t0 + next_cell_begin;
    const float cell_t1 = ray_t0 + fminf(t_closest, tMax);
    if (!lambda(cell, cell_t0, cell_t1)) return;
    next_cell_begin = t_closest;
    // a NaN ray steps no axis; end it instead of spinning (finite rays
    // always step at least one axis)
    bool stepped = false;
    if (t_next.x == t_closest) {
      t_next.x += t_step.x;
      cell.x += delta.x;
      stepped = true;
      if (cell.x == stop.x) return;
    }
    if (t_next.y == t_closest) {
      t_next.y += t_step.y;
      cell.y += delta.y;
      stepped = true;
      if (cell.y == stop.y) return;
    }
    if (t_next.z == t_closest) {
      t_next.z += t_step.z;
      cell.z += delta.z;
      stepped = true;
      if (cell.z == stop.z) return;
    }
    if (!stepped) return;
  }
}

inline int cell_id(const oracle_renderer* o, vec3i c) {
  return c.x + c.y * o->mc[0] + c.z * o->mc[0] * o->mc[1];
}

inline void record_hit(Hit& prd, float tWorld, vec4f s, int c) {
  prd.tHit = tWorld;
  prd.rgba = s;
  prd.rgba.w = 1.0f;
  prd.missed = false;
  prd.channel = c;
}

// deviceCode.cu:858-970 baseLineDTCH (-m 0, N-DDA)
void trace_ndda(const oracle_renderer* o, const GridRay& g, float prd_t1, Rng& rng, Hit& prd) {
  const int C = int(o->ch.size());
  const float unit = o->dt;
  int curMesh = 0;
  float tMax = 1e20f;
  uint32_t visit = 0;  // macrocells entered in this channel's pass: block offset
  auto lambda = [&](vec3i cellIdx, float t0, float t1) -> bool {
    prd.visits++;
    const int cid = cell_id(o, cellIdx);
    const uint32_t vis = visit++;
    const float majorant = o->maj[size_t(cid) * C + curMesh];
    if (majorant == 0.00f) return true;
    float t = t0;
    float u = rng.first_dist(vis, uint32_t(curMesh));
    while (true) {
      t = t - (flight_log(o, 1.0f - u) / majorant) * unit * g.worldToGridT;
      if (t >= t1) break;
      if (tMax < t) return false;
      const float tWorld = t * g.gridToWorldT;
      const vec3f xTexture = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
      if (tWorld >= prd_t1) {
        prd.rejections++;
        return false;
      }
      const float nullColThreshold = rng.pair_thr() * majorant;
      const float value = sample_volume(o, xTexture, curMesh);
      prd.samples++;
      if (!std::isnan(value)) {
        const vec4f s = transfer_function(o, value, curMesh);
        if (s.w > 0.0f && nullColThreshold < s.w) {
          record_hit(prd, tWorld, s, curMesh);
          tMax = fminf(tMax, t);
          return false;
        }
        prd.rejections++;
      }
      u = rng.pair_dist();
    }
    return true;
  };
  for (int i = 0; i < C; ++i) {
    curMesh = i;
    visit = 0;
    dda3(g.org, g.dir, 1e20f, o->mc, lambda);
  }
}

// deviceCode.cu:619-745 multiMajDTCH (-m 1): per cell one exponential clock
// per channel; the earliest clock fires, its sample is tested against that
// channel's majorant, and only that channel's clock is advanced. The draw
// order follows the reference: one draw per channel with a positive
// majorant on cell entry, then per event the next-step draw first and the
// threshold second (one Rng step per draw pair, see Rng).
void trace_multi(const oracle_renderer* o, const GridRay& g, float prd_t1, Rng& rng, Hit& prd) {
  const int C = int(o->ch.size());
  const float unit = o->dt;
  float majorants[kMaxChannels];
  float ts[kMaxChannels];
  uint32_t pairs[kMaxChannels];  // events of each channel's clock so far in the cell
  uint32_t visit = 0;            // macrocells entered so far: the streams' block offset
  auto lambda = [&](vec3i cellIdx, float t0, float t1) -> bool {
    prd.visits++;
    const int cid = cell_id(o, cellIdx);
    const uint32_t vis = visit++;
    float majorantSum = 0.0f;
    for (int i = 0; i < C; ++i) {
      majorants[i] = o->maj[size_t(cid) * C + i];
      majorantSum += majorants[i];
      ts[i] = t0;
      pairs[i] = 0;
    }
    if (majorantSum <= 0.0000001f) return true;
    // the reference draws one running stream in this order: a distance per
    // channel, then per event a distance (before the selection) and, for an
    // event inside the cell, its threshold
    for (int i = 0; i < C; ++i) {
      if (majorants[i] > 0.0f) {
        const float u = rng.seq ? rng.step() : rng.cell_word(vis, uint32_t(i));
        ts[i] = ts[i] - (flight_log(o, 1.0f - u) / majorants[i]) * unit * g.worldToGridT;
      } else {
        ts[i] = 1e20f;
      }
    }
    while (true) {
      float minT = ts[0];
      int sel = 0;
      const float rand_seq = rng.seq ? rng.step() : 0.0f;
      for (int i = 1; i < C; ++i) {
        if (ts[i] < minT && majorants[i] > 0.f) {
          sel = i;
          minT = ts[i];
        }
      }
      if (minT >= t1) break;
      const float tWorld = minT * g.gridToWorldT;
      const vec3f xTexture = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
      if (tWorld >= prd_t1) {
        prd.rejections++;
        return false;
      }
      float th, rand;
      if (rng.seq) {
        th = rng.thr();
        rand = rand_seq;
      } else {
        rng.vc_pair(vis, uint32_t(sel), pairs[sel]++, th, rand);
      }
      const float nullColThreshold = th * majorants[sel];
      const float value = sample_volume(o, xTexture, sel);
      prd.samples++;
      if (!std::isnan(value)) {
        const vec4f s = transfer_function(o, value, sel);
        if (s.w > 0.0f && nullColThreshold < s.w) {
          record_hit(prd, tWorld, s, sel);
          return false;
        }
      }
      prd.rejections++;
      ts[sel] = ts[sel] - (flight_log(o, 1.0f - rand) / majorants[sel]) * unit * g.worldToGridT;
    }
    return true;
  };
  dda3(g.org, g.dir, 1e20f, o->mc, lambda);
}

// deviceCode.cu:747-856 altMultiMajDTCH (-m 2, syncDDA)
void trace_sync(const oracle_renderer* o, const GridRay& g, float prd_t1, Rng& rng, Hit& prd) {
  const int C = int(o->ch.size());
  const float unit = o->dt;
  uint32_t visit = 0;  // macrocells entered so far: the stream's block offset
  auto lambda = [&](vec3i cellIdx, float t0, float t1) -> bool {
    prd.visits++;
    const int cid = cell_id(o, cellIdx);
    const uint32_t vis = visit++;
    float minT = t1;
    for (int c = 0; c < C; ++c) {
      const float majorant = o->maj[size_t(cid) * C + c];
      float t = t0;
      if (majorant == 0.0f) continue;
      float u = rng.first_dist(vis, uint32_t(c));
      while (true) {
        t = t - (flight_log(o, 1.0f - u) / majorant) * unit * g.worldToGridT;
        if (t >= minT) break;
        const float tWorld = t * g.gridToWorldT;
        const vec3f xTexture = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
        if (tWorld >= prd_t1) {
          prd.rejections++;
          if (c == C - 1) return false;
          break;
        }
        const float nullColThreshold = rng.pair_thr() * majorant;
        u = rng.pair_dist();
        const float value = sample_volume(o, xTexture, c);
        prd.samples++;
        if (!std::isnan(value)) {
          const vec4f s = transfer_function(o, value, c);
          if (s.w > 0.0f && nullColThreshold < s.w) {
            record_hit(prd, tWorld, s, c);
            minT = fminf(t, minT);
          }
        }
        prd.rejections++;
      }
    }
    return prd.missed;
  };
  dda3(g.org, g.dir, 1e20f, o->mc, lambda);
}

// deviceCode.cu:513-617 cummilativeDTCH (-m 3, cumulative majorant)
// The draws per visit: the first distance, then per step that lands inside
// the cell its threshold and the next distance (first_dist / pair_thr /
// pair_dist, channel 0's positions of the visit); in the faithful path the
// same calls take the running stream in the reference's order (distance,
// threshold, distance, ...).
void trace_cumulative(const oracle_renderer* o, const GridRay& g, float prd_t1, Rng& rng,
                      Hit& prd) {
  const int C = int(o->ch.size());
  const float unit = o->dt;
  uint32_t visit = 0;  // macrocells entered so far: the stream's block offset
  auto lambda = [&](vec3i cellIdx, float t0, float t1) -> bool {
    prd.visits++;
    const int cid = cell_id(o, cellIdx);
    const uint32_t vis = visit++;
    const float majorant = o->maj[size_t(cid)];
    if (majorant == 0.00f) return true;
    float t = t0;
    float u = rng.first_dist(vis, 0u);
    while (true) {
      t = t - (flight_log(o, 1.0f - u) / majorant) * unit * g.worldToGridT;
      if (t >= t1) break;
      const float tWorld = t * g.gridToWorldT;
      const vec3f xTexture = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
      if (tWorld >= prd_t1) {
        prd.rejections++;
        return false;
      }
      float nullColThreshold = rng.pair_thr() * majorant;
      u = rng.pair_dist();
      for (int c = 0; c < C; ++c) {
        const float value = sample_volume(o, xTexture, c);
        prd.samples++;
        if (std::isnan(value)) continue;
        const vec4f s = transfer_function(o, value, c);
        if (s.w > 0.0f && nullColThreshold < s.w) {
          record_hit(prd, tWorld, s, c);
          return false;
        }
        nullColThreshold -= s.w;
      }
      prd.rejections++;
    }
    return true;
  };
  dda3(g.org, g.dir, 1e20f, o->mc, lambda);
}

// deviceCode.cu:222-241 blendChannels with maxOpacityBlend (165-179, modes 4
// and 7), mixBlend (182-201, modes 5 and 8) and densityBlend (204-220, mode
// 6); (0,0,0,0) outside the unit cube.
vec4f blend_channels(const oracle_renderer* o, vec3f p) {
  if (p.x < 0.0f || p.x > 1.0f || p.y < 0.0f || p.y > 1.0f || p.z < 0.0f || p.z > 1.0f)
    return {0.f, 0.f, 0.f, 0.f};
  const int C = int(o->ch.size());
  if (o->mode == 6) {
    float densitySum = 0.0f;
    vec4f color = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < C; ++i) {
      const vec4f s = transfer_function(o, sample_volume(o, p, i), i);
      color = color + s.w * s;
      densitySum += s.w;
    }
    if (densitySum <= 0.0f) return {0.f, 0.f, 0.f, 0.f};
    return color / densitySum;
  }
  if (o->mode == 4 || o->mode == 7) {
    vec4f best = {0.f, 0.f, 0.f, 0.f};
    for (int c = 0; c < C; ++c) {
      const vec4f s = transfer_function(o, sample_volume(o, p, c), c);
      if (s.w > best.w) best = s;
    }
    return best;
  }
  vec3f color = {0.f, 0.f, 0.f};  // modes 5 and 8
  float maxOpacity = 0.0f;
  for (int c = 0; c < C; ++c) {
    const vec4f s = transfer_function(o, sample_volume(o, p, c), c);
    color = color + vec3f{s.x, s.y, s.z} * s.w;
    if (s.w > maxOpacity) maxOpacity = s.w;
  }
  if (maxOpacity <= 0.0f) return {0.f, 0.f, 0.f, 0.f};
  color = {color.x / maxOpacity, color.y / maxOpacity, color.z / maxOpacity};
  return {color.x, color.y, color.z, maxOpacity};
}

// deviceCode.cu:972-1065 blendDTCH (-m 4 MAX, -m 5 MIX): Woodcock tracking on
// the summed majorant with a blended sample. The reference leaves the
// payload's channel at 0 for these modes.
void trace_blend(const oracle_renderer* o, const GridRay& g, float prd_t1, Rng& rng, Hit& prd) {
  const float unit = o->dt;
  uint32_t visit = 0;  // macrocells entered so far: the stream's block offset
  auto lambda = [&](vec3i cellIdx, float t0, float t1) -> bool {
    prd.visits++;
    const uint32_t vis = visit++;
    const float majorant = o->maj[size_t(cell_id(o, cellIdx))];
    if (majorant == 0.00f) return true;
    float t = t0;
    // channel 0's positions of the visit, as -m 3; the faithful path draws
    // the threshold only for a positive blended opacity (the reference's
    // short-circuit), then the next distance
    float u = rng.first_dist(vis, 0u);
    while (true) {
      t = t - (flight_log(o, 1.0f - u) / majorant) * unit * g.worldToGridT;
      if (t >= t1) break;
      const float tWorld = t * g.gridToWorldT;
      const vec3f xTexture = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
      if (tWorld >= prd_t1) {
        prd.rejections++;
        return false;
      }
      const vec4f s = blend_channels(o, xTexture);
      prd.samples++;
      if (s.w > 0.0f && rng.pair_thr() * majorant < s.w) {
        record_hit(prd, tWorld, s, 0);
        return false;
      }
      u = rng.pair_dist();
      prd.rejections++;
    }
    return true;
  };
  dda3(g.org, g.dir, 1e20f, o->mc, lambda);
}

inline vec3f over3(vec3f Cin, vec3f Cx, float Ain, float Ax) {  // deviceCode.cu:44-47
  return Cin + Cx * Ax * (1.f - Ain);
}
inline float over1(float Ain, float Ax) { return Ain + (1.f - Ain) * Ax; }  // deviceCode.cu:49-52

// deviceCode.cu:1067-1157 rayMarcherCH (-m 6/7/8) with shadows off: front to
// back compositing of blended samples every dt from the root-box entry. The
// sample position trails t by one step and starts at the ray origin (the
// reference updates pos after sampling), which is reproduced.
void march_blend(const oracle_renderer* o, vec3f org, vec3f dir0, float prd_t0, float prd_t1,
                 Hit& prd) {
  const vec3f worldToUnit = 1.f / (o->bhi - o->blo);
  const float dt = o->dt;
  const vec3f dir = normalize(dir0);
  vec3f pos = org;
  float alpha = 0.0f;
  vec3f color = {0.f, 0.f, 0.f};
  const int C = int(o->ch.size());
  float highestContribution = 0.0f, highestContribDistance = 0.0f;
  for (float t = prd_t0; (t < prd_t1) && (alpha < 0.99f); t += dt) {
    const vec3f posTex = pos * worldToUnit;
    const vec4f b = blend_channels(o, posTex);
    prd.samples += uint64_t(C);
    color = over3(color, vec3f{b.x, b.y, b.z}, alpha, b.w);
    alpha = over1(alpha, b.w);
    if (o->enable_shadows && highestContribution < b.w) {
      highestContribution = b.w;
      highestContribDistance = t;
    }
    pos = org + dir * t;
  }
  // deviceCode.cu:1117-1154: one marched shadow ray from the point of
  // highest opacity towards the light. Its first sample sits one world unit
  // away from that point (posSh = orgSh + dirSh), as in the reference.
  vec3f shadow = {1.f, 1.f, 1.f};
  if (o->enable_shadows && highestContribution > 0.0f) {
    const vec3f orgSh = org + dir * highestContribDistance;
    const vec3f dirSh = -o->light_dir;
    float alphaSh = 0.f;
    vec3f posSh = orgSh + dirSh;
    const vec3f dirfrac = 1.f / dirSh;
    const float t1 = (o->blo.x - orgSh.x) * dirfrac.x, t2 = (o->bhi.x - orgSh.x) * dirfrac.x;
    const float t3 = (o->blo.y - orgSh.y) * dirfrac.y, t4 = (o->bhi.y - orgSh.y) * dirfrac.y;
    const float t5 = (o->blo.z - orgSh.z) * dirfrac.z, t6 = (o->bhi.z - orgSh.z) * dirfrac.z;
    float thit0 = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
    const float thit1 = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
    if ((thit1 >= 0) && (thit0 < thit1)) {
      thit0 = fmaxf(thit0, 0.00001f);
      if (length(dirSh) < 1e-5f) thit0 = thit1;
      for (float tSh = thit0; (tSh < thit1) && (alphaSh < 0.99f); tSh += dt) {
        const vec3f posTexSh = posSh * worldToUnit;
        const float opacitySh = blend_channels(o, posTexSh).w;
        alphaSh = over1(alphaSh, opacitySh);
        posSh = orgSh + dirSh * tSh;
      }
      prd.samples += uint64_t(C);
    }
    const float sh = (1.f - o->ambient) * (1.f - alphaSh) + o->ambient;
    shadow = {sh, sh, sh};
  }
  prd.missed = false;
  const vec3f cs = color * shadow;
  prd.rgba = {cs.x, cs.y, cs.z, alpha};
}

// deviceCode.cu:1159-1276 compRayMarcherCH (-m 9): each channel marched and
// composited on its own, then mixed by opacity. The reference does not reset
// the sample position between channels, so channel n > 0 starts where
// channel n-1 stopped; reproduced.
void march_comp(const oracle_renderer* o, vec3f org, vec3f dir0, float prd_t0, float prd_t1,
                Hit& prd) {
  const vec3f worldToUnit = 1.f / (o->bhi - o->blo);
  const float dt = o->dt;
  const vec3f dir = normalize(dir0);
  vec3f pos = org;
  const int C = int(o->ch.size());
  float alphas[kMaxChannels];
  vec3f colors[kMaxChannels];
  for (int i = 0; i < C; ++i) {
    alphas[i] = 0.0f;
    colors[i] = {0.f, 0.f, 0.f};
  }
  for (int n = 0; n < C; ++n) {
    for (float t = prd_t0; (t < prd_t1) && (alphas[n] < 0.99f); t += dt) {
      const vec3f posTex = pos * worldToUnit;
      const float v = sample_volume(o, posTex, n);
      vec4f sc = {0.f, 0.f, 0.f, 0.f};
      if (!std::isnan(v)) {
        sc = transfer_function(o, v, n);
        prd.samples++;
      }
      colors[n] = over3(colors[n], vec3f{sc.x, sc.y, sc.z}, alphas[n], sc.w);
      alphas[n] = over1(alphas[n], sc.w);
      pos = org + dir * t;
    }
  }
  float alphaSum = 0.0f;
  for (int i = 0; i < C; ++i) alphaSum += alphas[i];
  vec3f fin = {0.f, 0.f, 0.f};
  if (alphaSum > 0.0f)
    for (int i = 0; i < C; ++i) fin = fin + colors[i] * alphas[i] / alphaSum;
  prd.missed = false;
  prd.rgba = {fin.x, fin.y, fin.z, alphaSum / float(C)};
}

// deviceCode.cu:243-281 calculateGradient: central differences of the
// channel's TF opacity, one-sided at the faces of the unit cube.
vec3f calc_gradient(const oracle_renderer* o, vec3f p, float eps, int c) {
  float g[3];
  const float pc[3] = {p.x, p.y, p.z};
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
    const float v1 = transfer_function(o, sample_volume(o, {p1[0], p1[1], p1[2]}, c), c).w;
    const float v2 = transfer_function(o, sample_volume(o, {p2[0], p2[1], p2[2]}, c), c).w;
    g[i] = (v2 - v1) / div;
  }
  return {g[0], g[1], g[2]};
}

// deviceCode.cu:61-76 generateRay
inline void generate_ray(const oracle_renderer* o, float su, float sv, vec3f& org, vec3f& dir) {
  org = o->c_org;
  vec3f d = o->c_llc + su * o->c_hor + sv * o->c_ver;
  if (fabsf(d.x) < 1e-5f) d.x = 1e-5f;
  if (fabsf(d.y) < 1e-5f) d.y = 1e-5f;
  if (fabsf(d.z) < 1e-5f) d.z = 1e-5f;
  dir = normalize(d - org);
}

// deviceCode.cu:1676-1726 volumeIntersection against the root box
// (renderer.cpp:700-704: union of channel bounds); ray tmin = 0.
inline bool clip_root(const oracle_renderer* o, vec3f org, vec3f dir, float& t0, float& t1) {
  const vec3f lb = o->blo, rt = o->bhi;
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

inline uint32_t make_8bit(float f) {
  const float s = f * 256.f;
  // clamp in float first so that the conversion is always defined
  const float c = fminf(fmaxf(s, -1.f), 256.f);
  const int i = int(c);
  return uint32_t(std::min(255, std::max(0, i)));
}
inline uint32_t make_rgba(vec4f c) {
  return make_8bit(c.x) | (make_8bit(c.y) << 8) | (make_8bit(c.z) << 16) | (make_8bit(c.w) << 24);
}

// Root clip + the mode's closest-hit program for a Woodcock mode (0-5):
// deviceCode.cu:321 (primary) and 394 (shadow ray) traceRay.
bool trace_woodcock(const oracle_renderer* o, vec3f org, vec3f dir, Rng& rng, Hit& prd) {
  float t0, t1;
  if (!clip_root(o, org, dir, t0, t1)) return false;
  const GridRay g = grid_ray(o, org, dir);
  switch (o->mode) {
    case 0: trace_ndda(o, g, t1, rng, prd); break;
    case 1: trace_multi(o, g, t1, rng, prd); break;
    case 2: trace_sync(o, g, t1, rng, prd); break;
    case 3: trace_cumulative(o, g, t1, rng, prd); break;
    default: trace_blend(o, g, t1, rng, prd); break;
  }
  return true;
}

// deviceCode.cu:283-479 mainRG for one pixel and one frame (no surfaces,
// tMax = 1e20), with the gradient shading and shadow rays of 326-408 when
// enabled.
void render_pixel(oracle_renderer* o, int px, int py, uint64_t cnt[4]) {
  const size_t fbOfs = size_t(px) + size_t(o->W) * py;
  // deviceCode.cu:292 starts the sum at alpha 1: the stored alpha is 1 + 1/spp
  vec4f finalColor = {0.f, 0.f, 0.f, 1.f};
  const uint32_t pixel = uint32_t(fbOfs);
  for (int sampleID = 0; sampleID < o->spp; ++sampleID) {
    Rng rng(o->seed, pixel, o->frame_id, uint32_t(sampleID), 0u);
    rng.seq = o->faithful;
    float jx, jy;
    rng.jitter2(o->jitter2d, jx, jy);
    const float su = (float(px) + jx) / float(o->W);
    const float sv = (float(py) + jy) / float(o->H);
    vec3f org, dir;
    generate_ray(o, su, sv, org, dir);
    vec4f color = {0.f, 0.f, 0.f, 0.f};
    Hit prd;
    if (o->mode <= 5) {
      trace_woodcock(o, org, dir, rng, prd);
    } else {
      float t0, t1;
      if (clip_root(o, org, dir, t0, t1)) {
        if (o->mode <= 8)
          march_blend(o, org, dir, t0, t1, prd);
        else
          march_comp(o, org, dir, t0, t1, prd);
      }
    }
    if (!prd.missed) {
      // deviceCode.cu:322-433
      const vec3f albedo = {prd.rgba.x, prd.rgba.y, prd.rgba.z};
      const float transparency = prd.rgba.w;
      const Material& m = o->mat[prd.channel];
      const bool calculateShading = o->enable_gradient && (o->mode < 6) && length(albedo) > 0.001f;
      const bool calculateShadows = o->enable_shadows && (o->mode < 6);
      float diffuse = 1.0f, specular = 0.0f, shadow = 1.0f;
      if (calculateShading) {
        const vec3f worldToUnit = 1.f / (o->bhi - o->blo);
        const vec3f texSpacePos = (org + prd.tHit * dir) * worldToUnit;
        const vec3f gradient = calc_gradient(o, texSpacePos, m.gradient_step, prd.channel);
        if (length(gradient) > 1e-6f) {
          const vec3f N = normalize(-gradient);
          const vec3f L = normalize(o->light_dir);
          const vec3f V = normalize(-dir);
          const vec3f Hh = normalize(L + V);
          const float NdotL = fmaxf(dot(N, L), 0.0f);
          const float NdotH = fmaxf(dot(N, Hh), 0.0f);
          diffuse = m.diffuse * NdotL;
          if (NdotL > 0.0f && NdotH > 0.001f) {
            const float normalization = (m.shininess + 8.0f) / (8.0f * 3.14159265f);
            specular = m.specular * normalization * ppowf(NdotH, m.shininess);
          } else {
            specular = 0.0f;
          }
        } else {
          diffuse = m.diffuse;
          specular = 0.0f;
        }
      }
      if (calculateShadows) {
        // shadow ray from the hit towards the light, own RNG stream
        Rng srng(o->seed, pixel, o->frame_id, uint32_t(sampleID), 1u);
        srng.seq = o->faithful;
        Hit sp;
        trace_woodcock(o, org + prd.tHit * dir, -o->light_dir, srng, sp);
        const float sw = sp.missed ? 0.0f : sp.rgba.w;
        shadow = (1.f - o->ambient) * (1.f - sw);
        prd.samples += sp.samples;
        prd.rejections += sp.rejections;
        prd.visits += sp.visits;
      }
      const vec3f sc = m.spec;
      const vec3f mixedSpecularColor = albedo * 0.75f + sc * 0.25f;
      const vec3f ambientColor = albedo * 0.75f + sc * 0.25f;
      const vec3f diffuseContrib = diffuse * shadow * o->light_intensity * albedo;
      const vec3f specularContrib = specular * shadow * o->light_intensity * mixedSpecularColor;
      const vec3f ambientContrib = o->ambient * ambientColor;
      const vec3f shaded = diffuseContrib + specularContrib + ambientContrib;
      color = {shaded.x, shaded.y, shaded.z, transparency};
    }
    cnt[0] += prd.samples;
    cnt[1] += prd.visits;
    cnt[2] += prd.rejections;
    cnt[3] += 1;
    // over(color, vec4f(bg, 1)) (deviceCode.cu:44-59, 433)
    const float Ain = color.w, Ax = 1.f;
    const vec3f Cin = {color.x, color.y, color.z};
    const vec3f c = Cin + o->bg * Ax * (1.f - Ain);
    const float a = Ain + (1.f - Ain) * Ax;
    finalColor = finalColor + vec4f{c.x, c.y, c.z, a};
  }
  finalColor = finalColor / float(o->spp);
  // deviceCode.cu:463-469 running mean; the second moment is an addition
  const float k = float(o->accum_id);
  vec4f& acc = o->accum[fbOfs];
  vec4f& sq = o->m2[fbOfs];
  const vec4f c2 = {finalColor.x * finalColor.x, finalColor.y * finalColor.y,
                    finalColor.z * finalColor.z, finalColor.w * finalColor.w};
  if (o->accumulate) {
    acc = (finalColor + k * acc) / float(o->accum_id + 1);
    sq = (c2 + k * sq) / float(o->accum_id + 1);
  } else {
    acc = finalColor;
    sq = c2;
  }
}

}  // namespace

// ================================================================= C API ===
extern "C" {

oracle_renderer* oracle_create(void) { return new oracle_renderer(); }
void oracle_destroy(oracle_renderer* o) { delete o; }
const char* oracle_last_error(oracle_renderer* o) { return o->err.c_str(); }

// renderer.cpp:1031-1044 PushMesh; values converted as rawFile.cpp:100-123
int oracle_push_channel(oracle_renderer* o, const void* vox, int dtype, int nx, int ny, int nz) {
  if (o->inited) return fail(o, "push after init");
  if (int(o->ch.size()) >= kMaxChannels) return fail(o, "too many channels");
  if (dtype < 0 || dtype > 3) return fail(o, "bad dtype");
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(o, "bad dims");
  if (!o->ch.empty() && (o->ch[0].dims[0] != nx || o->ch[0].dims[1] != ny || o->ch[0].dims[2] != nz))
    return fail(o, "channel dims differ");
  Channel c;
  c.dtype = dtype;
  c.dims[0] = nx; c.dims[1] = ny; c.dims[2] = nz;
  const size_t n = size_t(nx) * ny * nz;
  const size_t bpv = dtype == ORACLE_U8 ? 1 : (dtype == ORACLE_F32 ? 4 : 2);
  c.raw.assign(static_cast<const uint8_t*>(vox), static_cast<const uint8_t*>(vox) + n * bpv);
  c.scale = dtype == ORACLE_U8 ? 255.f : (dtype == ORACLE_U16 ? 65535.f : 1.f);
  c.offset = 0.f;
  // bounds_.w starts at {0, 1} (rawFile.h:162) and is widened (rawFile.cpp:119-123)
  float lo = 0.f, hi = 1.f;
  for (size_t i = 0; i < n; ++i) {
    const float v = c.value_at(i);
    if (v > hi) hi = v;
    if (v < lo) lo = v;
  }
  c.vmin = lo;
  c.vmax = hi;
  o->ch.push_back(std::move(c));
  return 0;
}

// rawFile.cpp:228-237 reshapeBounds(remap): box [0,1]^3 -> [0, remap]
int oracle_set_bounds(oracle_renderer* o, const float remap[3]) {
  if (o->inited) return fail(o, "set_bounds after init");
  for (int i = 0; i < 3; ++i)
    if (!(remap[i] > 0.f) || !std::isfinite(remap[i])) return fail(o, "bad bounds");
  o->remap = {remap[0], remap[1], remap[2]};
  o->remap_set = true;
  return 0;
}

// renderer.cpp:324-942 Init (RAW path 638-903, camera 856-863, dt 935)
int oracle_init(oracle_renderer* o, int mode, const int mc[3], int auto_camera) {
  if (o->ch.empty()) return fail(o, "no channels");
  if (mode < 0 || mode > 9) return fail(o, "unsupported mode");
  o->mode = mode;
  // CalculateMCGridDims(4) (renderer.cpp:1342-1355) for any zero entry
  for (int i = 0; i < 3; ++i) {
    if (mc[i] < 0) return fail(o, "bad mc");
    if (mc[i] > 0) {
      o->mc[i] = mc[i];
    } else {
      unsigned acc = 0;
      for (const Channel& c : o->ch) acc += unsigned(float(c.dims[i]) / 4.f);
      unsigned avg = unsigned(float(acc) / float(o->ch.size()));
      o->mc[i] = avg < 1 ? 1 : int(avg);
    }
  }
  o->blo = {0.f, 0.f, 0.f};
  o->bhi = o->remap_set ? o->remap : vec3f{1.f, 1.f, 1.f};
  const int C = int(o->ch.size());
  o->dom_lo.resize(C);
  o->dom_hi.resize(C);
  o->tf.assign(C, TF());
  for (int c = 0; c < C; ++c) {
    o->dom_lo[c] = o->ch[c].vmin;
    o->dom_hi[c] = o->ch[c].vmax;
  }
  build_ranges(o);
  o->maj.assign(size_t(o->mc[0]) * o->mc[1] * o->mc[2] * (mode <= 2 ? C : 1), 0.f);
  if (auto_camera) {
    const vec3f center = {0.5f, 0.5f, 0.5f};
    const vec3f eye = {center.x, center.y, center.z + 2.5f};
    cam_set_orientation(o, eye, center, vec3f{0, 1, 0}, 45.0f);
  }
  o->inited = true;
  o->accum.assign(size_t(o->W) * o->H, vec4f{0, 0, 0, 0});
  o->m2.assign(size_t(o->W) * o->H, vec4f{0, 0, 0, 0});
  update_camera(o);
  o->frame_id = 0;
  o->accum_id = 0;
  oracle_reset_dt(o);
  return 0;
}

// renderer.cpp:1094-1193 SetXFColormap (alpha cubed), SetXFOpacityScale, SetXFRange
int oracle_set_tf(oracle_renderer* o, int ch, const float* rgba, int n_texels,
                  float opacity_scale, float xf_lo, float xf_hi) {
  if (!o->inited) return fail(o, "set_tf before init");
  if (ch < 0 || ch >= int(o->ch.size())) return fail(o, "bad channel");
  if (n_texels < 1 || n_texels > 4096) return fail(o, "bad texel count");
  if (!(xf_hi != xf_lo) || !std::isfinite(xf_lo) || !std::isfinite(xf_hi))
    return fail(o, "bad xf range");
  if (!(opacity_scale >= 0.f && opacity_scale <= 1000.f)) return fail(o, "bad opacity scale");
  for (int i = 0; i < n_texels; ++i)
    if (!(rgba[4 * i + 3] >= 0.f && rgba[4 * i + 3] <= 1.f)) return fail(o, "alpha outside [0,1]");
  TF& t = o->tf[ch];
  t.texels.resize(n_texels);
  for (int i = 0; i < n_texels; ++i)
    t.texels[i] = {rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], powf(rgba[4 * i + 3], 3.f)};
  t.opacity_scale = opacity_scale;
  t.xf_lo = xf_lo;
  t.xf_hi = xf_hi;
  t.set = true;
  o->accum_id = 0;
  if (tfs_ready(o)) build_majorants(o);
  return 0;
}

int oracle_set_camera(oracle_renderer* o, const float from[3], const float at[3],
                      const float up[3], float fovy_deg) {
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(from[i]) || !std::isfinite(at[i]) || !std::isfinite(up[i]))
      return fail(o, "non-finite camera");
  cam_set_orientation(o, {from[0], from[1], from[2]}, {at[0], at[1], at[2]},
                      {up[0], up[1], up[2]}, fovy_deg);
  if (o->inited) update_camera(o);
  return 0;
}

// renderer.cpp:1460-1466 Setdt
int oracle_set_dt(oracle_renderer* o, float dt) {
  o->dt = fmaxf(dt, 1e-4f);
  o->accum_id = 0;
  return 0;
}

// renderer.cpp:1444-1453 Resetdt (RAW): the loop assigns, so the last channel wins
int oracle_reset_dt(oracle_renderer* o) {
  if (o->ch.empty()) return fail(o, "no channels");
  float minVoxelSideLength = 3.402823466e38f;
  const vec3f span = o->bhi - o->blo;
  for (const Channel& c : o->ch)
    minVoxelSideLength = fminf(span.x / float(c.dims[0]),
                               fminf(span.y / float(c.dims[1]), span.z / float(c.dims[2])));
  return oracle_set_dt(o, minVoxelSideLength * 0.5f);
}

int oracle_set_frame(oracle_renderer* o, int w, int h, int spp, int accumulate, const float bg[3]) {
  if (w <= 0 || h <= 0 || spp <= 0) return fail(o, "bad frame");
  o->W = w;
  o->H = h;
  o->spp = spp;
  o->accumulate = accumulate != 0;
  o->bg = {bg[0], bg[1], bg[2]};
  o->accum.assign(size_t(w) * h, vec4f{0, 0, 0, 0});
  o->m2.assign(size_t(w) * h, vec4f{0, 0, 0, 0});
  if (o->inited) update_camera(o);
  o->accum_id = 0;
  return 0;
}

int oracle_set_seed(oracle_renderer* o, uint32_t seed) {
  o->seed = seed;
  o->accum_id = 0;
  return 0;
}

int oracle_set_faithful(oracle_renderer* o, int on) {
  o->faithful = on != 0;
  return 0;
}

int oracle_set_jitter_2d(oracle_renderer* o, int on) {
  o->jitter2d = on != 0;
  return 0;
}

int oracle_set_conservative_majorants(oracle_renderer* o, int on) {
  if (o->conservative_maj == (on != 0)) return 0;
  o->conservative_maj = on != 0;
  o->accum_id = 0;
  if (o->inited && tfs_ready(o)) build_majorants(o);
  return 0;
}

int oracle_set_threads(oracle_renderer* o, int n) {
  o->build_threads = n < 1 ? 1 : n;
  return 0;
}

int oracle_set_lighting(oracle_renderer* o, float light_intensity, float ambient) {
  o->light_intensity = light_intensity;
  o->ambient = ambient;
  o->accum_id = 0;
  return 0;
}

// SetLightDirection (renderer.cpp:1468-1474) and the enableShadows /
// enableGradientShading fields (renderer.h:121-124, the -l flag)
int oracle_set_light(oracle_renderer* o, const float dir[3], int shadows, int gradient) {
  o->light_dir = {dir[0], dir[1], dir[2]};
  o->enable_shadows = shadows != 0;
  o->enable_gradient = gradient != 0;
  o->accum_id = 0;
  return 0;
}

// SetMaterialDiffuse/Specular/Shininess/GradientStep/SpecularColor
// (renderer.cpp:1195-1248)
int oracle_set_material(oracle_renderer* o, int ch, float diffuse, float specular,
                        float shininess, float gradient_step, const float spec[3]) {
  if (ch < 0 || ch >= kMaxChannels) return fail(o, "bad channel");
  Material& m = o->mat[ch];
  m.diffuse = diffuse;
  m.specular = specular;
  m.shininess = shininess;
  m.gradient_step = gradient_step;
  m.spec = {spec[0], spec[1], spec[2]};
  o->accum_id = 0;
  return 0;
}

// renderer.cpp:944-973 Render: frame k of an accumulation uses accumID = k
int oracle_render(oracle_renderer* o, int n_frames, int y0, int y1, int nthreads) {
  if (!o->inited) return fail(o, "render before init");
  if (!tfs_ready(o)) return fail(o, "transfer functions not set");
  if (y1 <= 0 || y1 > o->H) y1 = o->H;
  if (y0 < 0) y0 = 0;
  for (int i = 0; i < 4; ++i) o->counters[i] = 0;
  for (int f = 0; f < n_frames; ++f) {
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) \
    reduction(+ : s0, s1, s2, s3)
#endif
    for (int py = y0; py < y1; ++py) {
      uint64_t cnt[4] = {0, 0, 0, 0};
      for (int px = 0; px < o->W; ++px) render_pixel(o, px, py, cnt);
      s0 += cnt[0]; s1 += cnt[1]; s2 += cnt[2]; s3 += cnt[3];
    }
    o->counters[0] += s0; o->counters[1] += s1; o->counters[2] += s2; o->counters[3] += s3;
    o->accum_id++;
    o->frame_id++;
  }
  return 0;
}

int oracle_reset_accumulation(oracle_renderer* o) {
  o->accum_id = 0;
  return 0;
}

int oracle_read_accum(oracle_renderer* o, float* rgba, float* m2) {
  const size_t n = size_t(o->W) * o->H;
  if (rgba) std::memcpy(rgba, o->accum.data(), n * 16);
  if (m2) std::memcpy(m2, o->m2.data(), n * 16);
  return 0;
}

int oracle_read_fb(oracle_renderer* o, uint32_t* out) {
  const size_t n = size_t(o->W) * o->H;
  for (size_t i = 0; i < n; ++i) out[i] = make_rgba(o->accum[i]);
  return 0;
}

int oracle_read_ranges(oracle_renderer* o, float* out) {
  if (!o->inited) return fail(o, "not initialized");
  std::memcpy(out, o->ranges.data(), o->ranges.size() * 4);
  return 0;
}

int oracle_read_majorants(oracle_renderer* o, float* out) {
  if (!o->inited) return fail(o, "not initialized");
  std::memcpy(out, o->maj.data(), o->maj.size() * 4);
  return 0;
}

int oracle_read_counters(oracle_renderer* o, uint64_t out[4]) {
  for (int i = 0; i < 4; ++i) out[i] = o->counters[i];
  return 0;
}

int oracle_get_info(oracle_renderer* o, int out[10]) {
  out[0] = o->mc[0]; out[1] = o->mc[1]; out[2] = o->mc[2];
  out[3] = o->ch.empty() ? 0 : o->ch[0].dims[0];
  out[4] = o->ch.empty() ? 0 : o->ch[0].dims[1];
  out[5] = o->ch.empty() ? 0 : o->ch[0].dims[2];
  out[6] = int(o->ch.size());
  out[7] = o->mode;
  out[8] = int(o->frame_id);
  out[9] = int(o->accum_id);
  return 0;
}

float oracle_get_dt(oracle_renderer* o) { return o->dt; }

int oracle_get_camera(oracle_renderer* o, float out[12]) {
  const vec3f v[4] = {o->c_org, o->c_llc, o->c_hor, o->c_ver};
  for (int i = 0; i < 4; ++i) {
    out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z;
  }
  return 0;
}

int oracle_get_domains(oracle_renderer* o, float* out) {
  for (size_t c = 0; c < o->ch.size(); ++c) {
    out[2 * c] = o->dom_lo[c];
    out[2 * c + 1] = o->dom_hi[c];
  }
  return 0;
}

int oracle_trace_cells(oracle_renderer* o, const float* org, const float* dir, int n,
                       int max_cells, int32_t* cells, float* t01, int32_t* counts) {
  if (!o->inited) return fail(o, "not initialized");
  for (int r = 0; r < n; ++r) {
    const GridRay g = grid_ray(o, {org[3 * r], org[3 * r + 1], org[3 * r + 2]},
                               {dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]});
    int k = 0;
    dda3(g.org, g.dir, 1e20f, o->mc, [&](vec3i c, float t0, float t1) -> bool {
      if (k < max_cells) {
        cells[size_t(r) * max_cells + k] = cell_id(o, c);
        t01[2 * (size_t(r) * max_cells + k)] = t0;
        t01[2 * (size_t(r) * max_cells + k) + 1] = t1;
      }
      ++k;
      return true;
    });
    for (int j = k; j < max_cells; ++j) {  // unused slots: cell -1, t = 0
      cells[size_t(r) * max_cells + j] = -1;
      t01[2 * (size_t(r) * max_cells + j)] = 0.f;
      t01[2 * (size_t(r) * max_cells + j) + 1] = 0.f;
    }
    counts[r] = k;
  }
  return 0;
}

int oracle_gen_rays(oracle_renderer* o, uint32_t frame, uint32_t sample, float* org, float* dir) {
  if (!o->inited) return fail(o, "not initialized");
  for (int py = 0; py < o->H; ++py)
    for (int px = 0; px < o->W; ++px) {
      const size_t i = size_t(px) + size_t(o->W) * py;
      Rng rng(o->seed, uint32_t(i), frame, sample, 0u);
      float jx, jy;
      rng.jitter2(o->jitter2d, jx, jy);
      const float su = (float(px) + jx) / float(o->W);
      const float sv = (float(py) + jy) / float(o->H);
      vec3f ro, rd;
      generate_ray(o, su, sv, ro, rd);
      org[3 * i] = ro.x; org[3 * i + 1] = ro.y; org[3 * i + 2] = ro.z;
      dir[3 * i] = rd.x; dir[3 * i + 1] = rd.y; dir[3 * i + 2] = rd.z;
    }
  return 0;
}

int oracle_clip_rays(oracle_renderer* o, const float* org, const float* dir, int n, float* t01,
                     int32_t* hit) {
  for (int r = 0; r < n; ++r) {
    float t0 = 0.f, t1 = 0.f;
    hit[r] = clip_root(o, {org[3 * r], org[3 * r + 1], org[3 * r + 2]},
                       {dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]}, t0, t1) ? 1 : 0;
    t01[2 * r] = t0;
    t01[2 * r + 1] = t1;
  }
  return 0;
}

void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], int rounds, uint32_t out[4]) {
  philox(ctr, key, out, rounds);
}
int oracle_philox_rounds(void) { return kPhiloxRounds; }

float oracle_logf(float x) { return plogf(x); }
float oracle_logf_poly(float x) { return plogf_poly(x); }
float oracle_expf(float x) { return pexpf(x); }
float oracle_powf(float x, float y) { return ppowf(x, y); }

float oracle_sample(oracle_renderer* o, int ch, const float p[3]) {
  return sample_volume(o, {p[0], p[1], p[2]}, ch);
}

void oracle_tf(oracle_renderer* o, int ch, float value, float out[4]) {
  const vec4f v = transfer_function(o, value, ch);
  out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
}

}  // extern "C"
