// C ABI of libmdwt (include/mdwt.h): the host side of the dtracker::Renderer
// RAW path (renderer.h:52-222, renderer.cpp) re-designed for one HIP device and
// one stream per handle. Host-side math that feeds the kernels (camera basis,
// dt, macrocell dims, projection windows) follows the reference's fp32
// expressions in order; the kernels do the rest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "batch_deal.hpp"
#include "mdwt.h"
#include "mdwt_internal.hpp"
#include "mdwt_types.hpp"

namespace {

struct F3 {
  float x, y, z;
};
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 normalize(F3 a) { return a * (1.f / std::sqrt(dot(a, a))); }
inline F3 cross(F3 a, F3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Float4 {
  float x, y, z, w;
};

thread_local std::string g_create_error;

}  // namespace

// renderer.h:121-142 defaults: shadows and gradient shading off, light
// from above, MaterialData per channel
inline mdwt::ShadeParams default_shade() {
  mdwt::ShadeParams s{};
  s.light_dir[0] = 0.0001f;
  s.light_dir[1] = -1.f;
  s.light_dir[2] = 0.0001f;
  for (int c = 0; c < mdwt::kMaxChannels; ++c) {
    s.diffuse[c] = 0.8f;
    s.specular[c] = 0.3f;
    s.shininess[c] = 8.0f;
    s.grad_step[c] = 0.01f;
    s.spec[c][0] = s.spec[c][1] = s.spec[c][2] = 1.f;
  }
  return s;
}

// Frame-kernel schedule of a new handle: the wave-coherent kernel for -m 0 to
// -m 5 (kernels_wave.hip), the thread-per-pixel kernel everywhere else.
constexpr int kDefaultSchedule = 1;

struct mdwt_renderer {
  int device = 0;
  std::string err;
  bool hip_ready = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // channels staged on the host until init (RawR keeps a host copy too)
  std::vector<std::vector<uint8_t>> staging;
  int dtype = -1;
  int dims[3] = {0, 0, 0};
  int C = 0;
  float vmin[MDWT_MAX_CHANNELS] = {}, vmax[MDWT_MAX_CHANNELS] = {};
  F3 remap{1.f, 1.f, 1.f};

  bool inited = false;
  int mode = 0;
  int mc[3] = {0, 0, 0};

  // transfer functions (TFData, renderer.h:39-50)
  std::vector<std::vector<Float4>> cm;  // alpha cubed
  float os[MDWT_MAX_CHANNELS], xlo[MDWT_MAX_CHANNELS], xhi[MDWT_MAX_CHANNELS];
  bool cm_set[MDWT_MAX_CHANNELS] = {};
  int tf_stride = 0;
  bool maj_valid = false;

  // camera (camera.h:46-142 state) and LaunchParams::camera
  F3 cam_pos{0, -1, 0}, fvx{1, 0, 0}, fvy{0, 1, 0}, fvz{0, 0, 1}, cam_up{0, 1, 0};
  float fovy = 60.f;
  mdwt::CameraParams cam{};

  // frame state
  int W = 1024, H = 1024, spp = 1;
  bool accumulate = true;
  float bg[3] = {0.f, 0.f, 0.f};
  float light = 1.0f, ambient = 0.1f;   // renderer.h:124-126
  mdwt::ShadeParams sh = default_shade();
  uint32_t seed = 0x5eed1234u;
  uint32_t frame_id = 0, accum_id = 0;
  float dt = 0.05f;
  int order_mode = 2;  // cost-sorted block order
  float last_kernel_ms = 0.f;

  // device memory
  uint8_t* d_vol = nullptr;
  uint64_t chan_stride = 0;
  uint8_t* d_bricks = nullptr;  // brick copy of d_vol, read by the samplers
  uint64_t brick_stride = 0;
  float* d_ranges = nullptr;
  float* d_maj = nullptr;
  Float4* d_tf = nullptr;
  size_t d_tf_texels = 0;
  Float4* d_accum = nullptr;
  Float4* d_m2 = nullptr;             // only while moments is on
  bool moments = false;                // second-moment buffers (mdwt_set_moments)
  size_t fb_pixels = 0;
  unsigned long long* d_counters = nullptr;
  uint32_t* d_order = nullptr;
  std::vector<uint32_t> order;  // host copy of d_order
  int n_blocks = 0;
  int* d_win = nullptr;
  uint32_t* d_wave_cost = nullptr;  // per-wave ticks of the last recorded frame
  bool aov_on = false;
  int aov_frames = 0;                // accumulate the buffers over the first frames only (0: all)
  Float4* d_aov = nullptr;          // 2 per pixel when aov_on
  bool cost_ready = false;
  bool need_resort = true;  // record costs and re-sort after the next frames

  // frame-kernel schedule (0: thread per pixel, 1: wave-coherent task queues)
  int schedule = kDefaultSchedule;
  bool jitter2d = false;  // y jitter from its own word (mdwt_set_jitter_2d)
  bool conservative_maj = false;  // mdwt_set_conservative_majorants
  int cu_count = 0;

  // workgroup pixel block (64: 8x8, 128: 16x8, 256: 16x16 threads/pixels) and
  // the Hilbert tile (pixels) of the block order
  int bw = 16, bh = 16, tile = 16;

  // batched views
  std::vector<mdwt::ViewParams> views;   // device accumulators in accum / m2
  mdwt::ViewParams* d_views = nullptr;
  std::vector<uint32_t> items;           // (view << 20) | (by << 10) | bx, issue order
  uint32_t* d_items = nullptr;
  unsigned* d_work = nullptr;            // [1] frame id, [2] accumulation id of launch form 2
  unsigned long long* d_item_ticks = nullptr;
  // launch form 2's item list: the items in the order the persistent grid
  // takes them (workgroup g takes positions g, g + G, ...), re-dealt by the
  // measured item costs after each call; pitems_measured when pitem_ticks
  // holds the last form-2 frame's stamps of this list (a host copy, so a
  // form-1 call in between does not change them)
  std::vector<uint32_t> pitems;
  uint32_t* d_pitems = nullptr;
  std::vector<unsigned long long> pitem_ticks;
  bool pitems_measured = false;
  uint32_t batch_frame = 0, batch_accum = 0;
  int batch_grid = 0, batch_grid_key = -1;  // resident grid of launch form 2
  hipGraph_t graph = nullptr;               // launch form 2: one captured frame
  hipGraphExec_t graph_exec = nullptr;
  bool graph_valid = false;
  mdwt::RenderParams graph_params;  // what the captured graph launches
  int graph_grid = 0;
  // per-view launches (the batch path in use): each view's block order and
  // block count, events around each view's launch, the last frame's times
  std::vector<uint32_t*> view_order;
  std::vector<int> view_blocks;
  std::vector<hipEvent_t> view_ev, view_ev_end;  // around each view's launch
  std::vector<hipEvent_t> join_ev;               // end of each batch stream's work
  std::vector<hipStream_t> batch_streams;        // views run round-robin on these
  std::vector<float> view_us;
  // 0: a launch per view over the batch streams, 1: one launch for all, 2: one
  // persistent launch for all, replayed from a captured HIP graph
  int batch_launch = 0;
};

namespace {

void free_batch(mdwt_renderer* h);
int reorder(mdwt_renderer* h);

int fail(mdwt_renderer* h, int code, const std::string& msg) {
  h->err = msg;
  return code;
}

// ResetAccumulation (renderer.cpp:1488-1495); every change that restarts the
// accumulation may also change where the frame's work lies, so the
// cost-sorted block order is re-measured after it.
void reset_acc(mdwt_renderer* h) {
  h->accum_id = 0;
  h->need_resort = true;
}

#define HIPCHK(h, expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((h), MDWT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t bytes_per_voxel(int dt) { return dt == MDWT_U8 ? 1 : (dt == MDWT_F32 ? 4 : 2); }

float half_to_float(uint16_t h) {
  const uint32_t s = uint32_t(h >> 15) << 31;
  uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff, bits;
  if (e == 0) {
    if (m == 0) {
      bits = s;
    } else {
      int ee = -1;
      do {
        m <<= 1;
        ++ee;
      } while ((m & 0x400) == 0);
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

float value_scale(int dt) { return dt == MDWT_U8 ? 255.f : (dt == MDWT_U16 ? 65535.f : 1.f); }

// camera.h:83-110 setOrientation with forceUpFrame (camera.h:72-80)
void set_orientation(mdwt_renderer* h, F3 origin, F3 interest, F3 up, float fovy) {
  h->fovy = fovy;
  h->cam_pos = origin;
  h->cam_up = up;
  const bool same = origin.x == interest.x && origin.y == interest.y && origin.z == interest.z;
  h->fvz = same ? F3{0, 0, 1} : -normalize(interest - origin);
  h->fvx = cross(up, h->fvz);
  if (dot(h->fvx, h->fvx) < 1e-8f)
    h->fvx = F3{0, 1, 0};
  else
    h->fvx = normalize(h->fvx);
  h->fvy = normalize(cross(h->fvz, h->fvx));
  if (std::fabs(dot(h->fvz, h->cam_up)) < 1e-6f) return;
  h->fvx = normalize(cross(h->cam_up, h->fvz));
  h->fvy = normalize(cross(h->fvz, h->fvx));
}

// renderer.cpp:1058-1092 UpdateCamera (+ ResetAccumulation)
void update_camera(mdwt_renderer* h) {
  const F3 lookFrom = h->cam_pos;
  const F3 lookAt = h->cam_pos - h->fvz;
  const F3 lookUp = h->fvy;
  const float cosFovy = std::cos(h->fovy * float(M_PI / 180.0));
  const float vfov = std::acos(cosFovy) * float(180.0 / M_PI);
  const F3 vup = lookUp;
  const float aspect = float(h->W) / float(h->H);
  const float theta = vfov * float(M_PI) / 180.0f;
  const float half_height = std::tan(theta / 2.0f);
  const float half_width = aspect * half_height;
  const float focusDist = 10.f;
  const F3 origin = lookFrom;
  const F3 w = normalize(lookFrom - lookAt);
  const F3 u = normalize(cross(vup, w));
  const F3 v = cross(w, u);
  const F3 llc = origin - half_width * focusDist * u - half_height * focusDist * v - focusDist * w;
  const F3 hor = 2.0f * half_width * focusDist * u;
  const F3 ver = 2.0f * half_height * focusDist * v;
  const F3 all[4] = {origin, llc, hor, ver};
  float* dst[4] = {h->cam.org, h->cam.llc, h->cam.hor, h->cam.ver};
  for (int i = 0; i < 4; ++i) {
    dst[i][0] = all[i].x;
    dst[i][1] = all[i].y;
    dst[i][2] = all[i].z;
  }
  reset_acc(h);
}

// renderer.cu:387-391 project
int project(float f, float lo, float hi, int dim) {
  const float q = float(dim) * (f - lo) / (hi - lo);
  const int i = int(q);
  return std::max(0, std::min(dim - 1, i));
}

// Per-axis voxel windows of the range grid. Voxel i's widened box along this
// axis (rasterElements RAW, renderer.cu:853-890: own box extended by the
// lower corners of its in-range neighbours) projects to macrocells
// [project(lower), min(project(upper), M-1)] (rasterBox float2,
// renderer.cu:487-519). The voxels covering macrocell m form an interval
// [a, b]; the cell's range is the min/max over [a-1, b+1] (each voxel brings
// its 3x3x3 neighbourhood). Returns false if a window is not an interval.
bool axis_windows(int n, int m, float wlo, float whi, std::vector<int>& out) {
  const float boxLen = (whi - wlo) / float(n);
  std::vector<int> clo(n), chi(n);
  std::vector<char> valid(n);
  for (int i = 0; i < n; ++i) {
    const float vxl = wlo + float(i) * boxLen;
    float lower = vxl, upper = vxl + boxLen;
    for (int d = -1; d <= 1; ++d) {
      const int q = i + d;
      if (q < 0 || q >= n) continue;
      const float nl = wlo + float(q) * boxLen;
      lower = std::fmin(lower, nl);
      upper = std::fmax(upper, nl);
    }
    valid[i] = lower < upper;
    clo[i] = project(lower, wlo, whi, m);
    chi[i] = std::min(project(upper, wlo, whi, m), m - 1);
  }
  out.assign(2 * size_t(m), 0);
  for (int c = 0; c < m; ++c) {
    int a = -1, b = -2;
    bool ended = false;
    for (int i = 0; i < n; ++i) {
      const bool in = valid[i] && clo[i] <= c && c <= chi[i];
      if (in) {
        if (ended) return false;
        if (a < 0) a = i;
        b = i;
      } else if (a >= 0) {
        ended = true;
      }
    }
    if (a < 0) {
      out[2 * c] = 1;  // empty
      out[2 * c + 1] = 0;
    } else {
      out[2 * c] = std::max(a - 1, 0);
      out[2 * c + 1] = std::min(b + 1, n - 1);
    }
  }
  return true;
}

// 2-D Hilbert curve, iterative quadrant rotation (Hamilton; the textbook
// d -> (x, y) walk), reported as (y, x) so that it traces the same curve as
// the reference's hilbert_i2c with nDims = 2 (hilbert.h:128-170).
void hilbert_d2xy(int bits, uint32_t d, uint32_t& ox, uint32_t& oy) {
  uint32_t x = 0, y = 0, t = d;
  const uint32_t n = 1u << bits;
  for (uint32_t s = 1; s < n; s <<= 1) {
    const uint32_t rx = 1u & (t >> 1);
    const uint32_t ry = 1u & (t ^ rx);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      std::swap(x, y);
    }
    x += s * rx;
    y += s * ry;
    t >>= 2;
  }
  ox = y;
  oy = x;
}

void free_frame(mdwt_renderer* h) {
  if (h->d_aov) (void)hipFree(h->d_aov);
  h->d_aov = nullptr;
  if (h->d_accum) (void)hipFree(h->d_accum);
  if (h->d_m2) (void)hipFree(h->d_m2);
  if (h->d_order) (void)hipFree(h->d_order);
  if (h->d_wave_cost) (void)hipFree(h->d_wave_cost);
  h->d_accum = h->d_m2 = nullptr;
  h->d_order = nullptr;
  h->d_wave_cost = nullptr;
  h->cost_ready = false;
  h->fb_pixels = 0;
}

// Cost-sorted block order (longest processing time first): blocks sorted by
// the summed wall-clock ticks of their waves in the last recorded frame,
// heaviest first. Workgroup i runs on XCD i % 8, so the sorted list is dealt
// round-robin over the XCDs and the slowest blocks start first instead of
// trailing at the end of the frame.
int resort_blocks(mdwt_renderer* h) {
  const size_t nb = size_t(h->n_blocks);
  const size_t wpb = size_t(h->bw * h->bh / 64);  // waves per block
  std::vector<uint32_t> ticks(wpb * nb);
  HIPCHK(h, hipMemcpyAsync(ticks.data(), h->d_wave_cost, ticks.size() * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<std::pair<uint64_t, uint32_t>> cost(nb);
  for (size_t s = 0; s < nb; ++s) {
    uint64_t c = 0;
    for (size_t w = 0; w < wpb; ++w) c += ticks[wpb * s + w];
    cost[s] = {c, h->order[s]};
  }
  std::stable_sort(cost.begin(), cost.end(),
                   [](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) {
                     return a.first > b.first;
                   });
  for (size_t s = 0; s < nb; ++s) h->order[s] = cost[s].second;
  HIPCHK(h, hipMemcpyAsync(h->d_order, h->order.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice,
                           h->stream));
  h->cost_ready = false;
  return MDWT_OK;
}

// Block order of one image (nbx x nby blocks of bw x bh pixels): row-major,
// or a Hilbert curve over tile x tile-pixel super-tiles (blocks row-major
// inside a super-tile). Entries are bx | by << 16.
std::vector<uint32_t> block_order(int nbx, int nby, int bw, int bh, int tile, bool hilbert) {
  std::vector<uint32_t> order;
  order.reserve(size_t(nbx) * nby);
  if (!hilbert) {
    for (int y = 0; y < nby; ++y)
      for (int x = 0; x < nbx; ++x) order.push_back(uint32_t(x) | (uint32_t(y) << 16));
    return order;
  }
  const int tx = std::max(1, tile / bw), ty = std::max(1, tile / bh);
  const int sx = (nbx + tx - 1) / tx, sy = (nby + ty - 1) / ty;
  int bits = 0;
  while ((1 << bits) < std::max(sx, sy)) ++bits;
  const uint64_t total = uint64_t(1) << (2 * bits);
  for (uint64_t d = 0; d < total; ++d) {
    uint32_t x, y;
    hilbert_d2xy(bits, uint32_t(d), x, y);
    if (int(x) >= sx || int(y) >= sy) continue;
    for (int j = 0; j < ty; ++j)
      for (int i = 0; i < tx; ++i) {
        const int bx = int(x) * tx + i, by = int(y) * ty + j;
        if (bx < nbx && by < nby) order.push_back(uint32_t(bx) | (uint32_t(by) << 16));
      }
  }
  return order;
}

// frame buffers and the workgroup -> pixel-block order of the frame kernel
int alloc_frame(mdwt_renderer* h) {
  free_frame(h);
  const size_t n = size_t(h->W) * h->H;
  HIPCHK(h, hipMalloc(&h->d_accum, n * sizeof(Float4)));
  HIPCHK(h, hipMemsetAsync(h->d_accum, 0, n * sizeof(Float4), h->stream));
  if (h->moments) {
    HIPCHK(h, hipMalloc(&h->d_m2, n * sizeof(Float4)));
    HIPCHK(h, hipMemsetAsync(h->d_m2, 0, n * sizeof(Float4), h->stream));
  }
  h->fb_pixels = n;
  if (h->aov_on) {
    const size_t na = size_t(mdwt::aov_groups(h->C)) * n;
    HIPCHK(h, hipMalloc(&h->d_aov, na * sizeof(Float4)));
    HIPCHK(h, hipMemsetAsync(h->d_aov, 0, na * sizeof(Float4), h->stream));
  }
  const int bw = h->bw, bh = h->bh;
  const int nbx = (h->W + bw - 1) / bw, nby = (h->H + bh - 1) / bh;
  std::vector<uint32_t> order = block_order(nbx, nby, bw, bh, h->tile, h->order_mode >= 1);
  h->n_blocks = int(order.size());
  HIPCHK(h, hipMalloc(&h->d_order, order.size() * sizeof(uint32_t)));
  HIPCHK(h, hipMemcpyAsync(h->d_order, order.data(), order.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMalloc(&h->d_wave_cost, 4 * order.size() * sizeof(uint32_t)));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->order = std::move(order);
  h->cost_ready = false;
  h->need_resort = true;
  return MDWT_OK;
}

bool all_tf_set(const mdwt_renderer* h) {
  for (int c = 0; c < h->C; ++c)
    if (!h->cm_set[c]) return false;
  return true;
}

mdwt::TfParams tf_params(const mdwt_renderer* h) {
  mdwt::TfParams t{};
  t.tab = h->d_tf;
  t.stride = h->tf_stride;
  for (int c = 0; c < h->C; ++c) {
    t.n[c] = int(h->cm[c].size());
    t.opacity[c] = h->os[c];
    t.xf_lo[c] = h->xlo[c];
    t.xf_hi[c] = h->xhi[c];
    t.dom_lo[c] = h->vmin[c];
    t.dom_hi[c] = h->vmax[c];
  }
  return t;
}

// RecalculateDensityRanges (renderer.cu:206-274): upload the TF texels and
// rebuild the majorant grid for the current mode.
int rebuild_majorants(mdwt_renderer* h) {
  h->maj_valid = false;
  if (!h->inited || !all_tf_set(h)) return MDWT_OK;
  int stride = 0;
  for (int c = 0; c < h->C; ++c) stride = std::max(stride, int(h->cm[c].size()));
  stride = (stride + 3) & ~3;
  if (size_t(h->C) * stride > size_t(mdwt::kMaxTexelsLds))
    return fail(h, MDWT_ERR_LIMIT, "transfer functions exceed 160 KiB of LDS (channels * texels)");
  const size_t texels = size_t(h->C) * stride;
  if (texels > h->d_tf_texels) {
    if (h->d_tf) (void)hipFree(h->d_tf);
    h->d_tf = nullptr;
    HIPCHK(h, hipMalloc(&h->d_tf, texels * sizeof(Float4)));
    h->d_tf_texels = texels;
  }
  h->tf_stride = stride;
  std::vector<Float4> tab(texels, Float4{0, 0, 0, 0});
  for (int c = 0; c < h->C; ++c)
    std::copy(h->cm[c].begin(), h->cm[c].end(), tab.begin() + size_t(c) * stride);
  HIPCHK(h, hipMemcpyAsync(h->d_tf, tab.data(), texels * sizeof(Float4), hipMemcpyHostToDevice,
                           h->stream));
  mdwt::MajorantParams p{};
  p.ranges = h->d_ranges;
  p.tf = tf_params(h);
  p.channels = h->C;
  p.n_cells = h->mc[0] * h->mc[1] * h->mc[2];
  p.mode = h->mode;
  p.conservative = h->conservative_maj ? 1 : 0;
  p.maj = h->d_maj;
  HIPCHK(h, mdwt::launch_majorants(p, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->maj_valid = true;
  return MDWT_OK;
}

mdwt::VolumeParams volume_params(const mdwt_renderer* h) {
  mdwt::VolumeParams v{};
  v.vol = h->d_vol;
  v.chan_stride = h->chan_stride;
  v.bricks = h->d_bricks;
  v.brick_stride = h->brick_stride;
  v.nbx = (h->dims[0] + 3) / 4;
  v.nby = (h->dims[1] + 3) / 4;
  v.nx = h->dims[0];
  v.ny = h->dims[1];
  v.nz = h->dims[2];
  v.channels = h->C;
  v.scale = value_scale(h->dtype);
  v.offset = 0.f;
  v.blo[0] = v.blo[1] = v.blo[2] = 0.f;
  v.bhi[0] = h->remap.x;
  v.bhi[1] = h->remap.y;
  v.bhi[2] = h->remap.z;
  return v;
}

// Allocate and fill the brick copy of the volume.
int make_bricks(mdwt_renderer* h) {
  if (h->d_bricks) return MDWT_OK;
  const size_t per = size_t((h->dims[0] + 3) / 4) * size_t((h->dims[1] + 3) / 4) *
                     size_t((h->dims[2] + 3) / 4) * 64u * bytes_per_voxel(h->dtype);
  h->brick_stride = (per + 255) & ~size_t(255);
  if (hipMalloc(&h->d_bricks, h->brick_stride * h->C) != hipSuccess) {
    h->d_bricks = nullptr;
    return fail(h, MDWT_ERR_OOM, "device allocation of the volume bricks failed");
  }
  mdwt::VolumeParams v = volume_params(h);
  v.bricks = h->d_bricks;
  HIPCHK(h, mdwt::launch_bricks(v, h->dtype, h->stream));
  return MDWT_OK;
}

// The wave-coherent schedule covers the tracking modes -m 0 to -m 5; the ray
// marchers, shadow rays, gradient shading and the
// frames that write auxiliary outputs (the first aov_frames of an
// accumulation) run the thread-per-pixel kernel (the launcher checks the same
// conditions). Both give the same image bits, so an accumulation may switch.
int eff_schedule(const mdwt_renderer* h) {
  if (h->sh.shadows || h->sh.gradient) return 0;
  if (h->aov_on && (h->aov_frames == 0 || h->accum_id < uint32_t(h->aov_frames))) return 0;
  return (h->mode >= 0 && h->mode <= 5) ? h->schedule : 0;
}

mdwt::RenderParams render_params(const mdwt_renderer* h) {
  mdwt::RenderParams p;
  std::memset(&p, 0, sizeof(p));  // padding too: batched launches compare params bytewise
  p.vol = volume_params(h);
  p.tf = tf_params(h);
  p.cam = h->cam;
  for (int i = 0; i < 3; ++i) p.mc[i] = h->mc[i];
  p.maj = h->d_maj;
  p.dt = h->dt;
  p.fr.width = h->W;
  p.fr.height = h->H;
  p.fr.spp = h->spp;
  p.fr.accumulate = h->accumulate ? 1 : 0;
  for (int i = 0; i < 3; ++i) p.fr.bg[i] = h->bg[i];
  p.fr.sh = h->sh;
  p.fr.light_intensity = h->light;
  p.fr.ambient = h->ambient;
  p.fr.seed = h->seed;
  p.fr.jitter2d = h->jitter2d ? 1 : 0;
  p.fr.frame_id = h->frame_id;
  p.fr.accum_id = h->accum_id;
  p.fr.accum = h->d_accum;
  p.fr.m2 = h->d_m2;
  p.fr.counters = h->d_counters;
  p.fr.block_order = h->d_order;
  p.fr.n_blocks = h->n_blocks;
  p.fr.schedule = eff_schedule(h);
  p.fr.wave_cost = (h->order_mode == 2 && h->need_resort) ? h->d_wave_cost : nullptr;
  p.fr.bw = h->bw;
  p.fr.bh = h->bh;
  p.fr.views = nullptr;
  p.fr.items = nullptr;
  p.fr.n_items = 0;
  p.fr.work = nullptr;
  p.fr.item_ticks = nullptr;
  const bool aov_now = h->aov_frames == 0 || h->accum_id < uint32_t(h->aov_frames);
  p.fr.aov = aov_now ? h->d_aov : nullptr;
  return p;
}

// One frame launch (the counting pass always runs the thread-per-pixel kernel).
int launch_frame(mdwt_renderer* h, bool count) {
  mdwt::RenderParams p = render_params(h);
  if (count) p.fr.wave_cost = nullptr;
  HIPCHK(h, mdwt::launch_render(p, h->mode, h->dtype, count, h->stream));
  if (p.fr.wave_cost) h->cost_ready = true;
  return MDWT_OK;
}

// (Re)allocates the value-range and majorant grids for h->mc and runs the
// range kernel over the resident volume (renderer.cu:1116-1179).
int build_range_grid(mdwt_renderer* h, const std::vector<int>& wx, const std::vector<int>& wy,
                     const std::vector<int>& wz) {
  const size_t n_cells = size_t(h->mc[0]) * h->mc[1] * h->mc[2];
  void* old[] = {h->d_win, h->d_ranges, h->d_maj};
  for (void* q : old)
    if (q) (void)hipFree(q);
  h->d_win = nullptr;
  h->d_ranges = nullptr;
  h->d_maj = nullptr;
  std::vector<int> win;
  win.insert(win.end(), wx.begin(), wx.end());
  win.insert(win.end(), wy.begin(), wy.end());
  win.insert(win.end(), wz.begin(), wz.end());
  HIPCHK(h, hipMalloc(&h->d_win, win.size() * sizeof(int)));
  HIPCHK(h, hipMemcpyAsync(h->d_win, win.data(), win.size() * sizeof(int), hipMemcpyHostToDevice,
                           h->stream));
  HIPCHK(h, hipMalloc(&h->d_ranges, n_cells * h->C * 2 * sizeof(float)));
  HIPCHK(h, hipMalloc(&h->d_maj, n_cells * h->C * sizeof(float)));
  HIPCHK(h, hipMemsetAsync(h->d_maj, 0, n_cells * h->C * sizeof(float), h->stream));
  mdwt::RangeParams rp{};
  rp.vol = h->d_vol;
  rp.chan_stride = h->chan_stride;
  rp.nx = h->dims[0];
  rp.ny = h->dims[1];
  rp.nz = h->dims[2];
  rp.channels = h->C;
  rp.scale = value_scale(h->dtype);
  rp.offset = 0.f;
  for (int i = 0; i < 3; ++i) rp.mc[i] = h->mc[i];
  rp.win = h->d_win;
  rp.ranges = h->d_ranges;
  HIPCHK(h, mdwt::launch_ranges(rp, h->dtype, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->maj_valid = false;
  return MDWT_OK;
}

int ready_to_render(mdwt_renderer* h) {
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (!all_tf_set(h) || !h->maj_valid)
    return fail(h, MDWT_ERR_STATE, "transfer functions not set for every channel");
  return MDWT_OK;
}

bool finite3(const float* v) {
  return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

}  // namespace

extern "C" {

int mdwt_create(mdwt_renderer** out, int hip_device) {
  if (!out) {
    g_create_error = "null output pointer";
    return MDWT_ERR_ARG;
  }
  *out = nullptr;
  if (hip_device < 0) {
    g_create_error = "negative device id";
    return MDWT_ERR_ARG;
  }
  mdwt_renderer* h = new (std::nothrow) mdwt_renderer();
  if (!h) {
    g_create_error = "out of host memory";
    return MDWT_ERR_OOM;
  }
  h->device = hip_device;
  for (int c = 0; c < MDWT_MAX_CHANNELS; ++c) {
    h->os[c] = 1.0f;  // TFData defaults (renderer.h:45-49)
    h->xlo[c] = 0.f;
    h->xhi[c] = 1.f;
  }
  *out = h;
  return MDWT_OK;
}

int mdwt_destroy(mdwt_renderer* h) {
  if (!h) return MDWT_ERR_ARG;
  if (h->hip_ready) {
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_frame(h);
    free_batch(h);
    void* ptrs[] = {h->d_vol, h->d_bricks, h->d_ranges, h->d_maj, h->d_tf, h->d_counters, h->d_win};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
  }
  delete h;
  return MDWT_OK;
}

const char* mdwt_last_error(const mdwt_renderer* h) {
  return h ? h->err.c_str() : g_create_error.c_str();
}

int mdwt_push_channel(mdwt_renderer* h, const void* voxels, int dtype, int nx, int ny, int nz) {
  if (!h) return MDWT_ERR_ARG;
  if (h->inited) return fail(h, MDWT_ERR_STATE, "channels must be pushed before init");
  if (h->C >= MDWT_MAX_CHANNELS)
    return fail(h, MDWT_ERR_LIMIT, "at most 12 channels (MAX_CHANNELS)");
  if (!voxels) return fail(h, MDWT_ERR_ARG, "null voxel pointer");
  if (dtype < MDWT_U8 || dtype > MDWT_F32) return fail(h, MDWT_ERR_ARG, "unknown dtype");
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail(h, MDWT_ERR_ARG, "dims must be positive");
  const uint64_t nvox = uint64_t(nx) * uint64_t(ny) * uint64_t(nz);
  if (nvox > 0xFFFFFFFFull / 2)
    return fail(h, MDWT_ERR_LIMIT, "a channel must hold fewer than 2^31 voxels");
  // the samplers' brick copy: 32-bit voxel indices, 24-bit products by the
  // bricks per slice
  const uint64_t bricks_xy = uint64_t((nx + 3) / 4) * uint64_t((ny + 3) / 4);
  if (bricks_xy >= (uint64_t(1) << 24) || bricks_xy * uint64_t((nz + 3) / 4) * 64u > 0xFFFFFFFFull)
    return fail(h, MDWT_ERR_LIMIT, "volume too large for the brick layout (at most 2^24 bricks per slice)");
  if (h->C > 0 && (nx != h->dims[0] || ny != h->dims[1] || nz != h->dims[2]))
    return fail(h, MDWT_ERR_ARG, "all channels must share dims");
  if (h->C > 0 && dtype != h->dtype) return fail(h, MDWT_ERR_ARG, "all channels must share dtype");
  const size_t bytes = size_t(nvox) * bytes_per_voxel(dtype);
  std::vector<uint8_t> buf;
  try {
    buf.assign(static_cast<const uint8_t*>(voxels), static_cast<const uint8_t*>(voxels) + bytes);
  } catch (...) {
    return fail(h, MDWT_ERR_OOM, "out of host memory staging the channel");
  }
  // value domain: bounds_.w starts at {0, 1} and widens (rawFile.h:162, rawFile.cpp:119-123)
  const float sc = value_scale(dtype);
  float lo = 0.f, hi = 1.f;
  for (size_t i = 0; i < size_t(nvox); ++i) {
    float r;
    if (dtype == MDWT_U8) {
      r = float(buf[i]);
    } else if (dtype == MDWT_U16) {
      uint16_t v;
      std::memcpy(&v, &buf[2 * i], 2);
      r = float(v);
    } else if (dtype == MDWT_F16) {
      uint16_t v;
      std::memcpy(&v, &buf[2 * i], 2);
      r = half_to_float(v);
    } else {
      std::memcpy(&r, &buf[4 * i], 4);
      if (!std::isfinite(r)) return fail(h, MDWT_ERR_ARG, "non-finite voxel value");
    }
    const float v = r * sc + 0.f;
    if (v > hi) hi = v;
    if (v < lo) lo = v;
  }
  if (dtype == MDWT_F16)
    for (size_t i = 0; i < size_t(nvox); ++i) {
      uint16_t v;
      std::memcpy(&v, &buf[2 * i], 2);
      if ((v & 0x7c00) == 0x7c00) return fail(h, MDWT_ERR_ARG, "non-finite voxel value");
    }
  h->vmin[h->C] = lo;
  h->vmax[h->C] = hi;
  h->staging.push_back(std::move(buf));
  h->dtype = dtype;
  h->dims[0] = nx;
  h->dims[1] = ny;
  h->dims[2] = nz;
  h->C += 1;
  return MDWT_OK;
}

int mdwt_set_bounds(mdwt_renderer* h, const float remap[3]) {
  if (!h) return MDWT_ERR_ARG;
  if (h->inited) return fail(h, MDWT_ERR_STATE, "bounds must be set before init");
  if (!remap || !finite3(remap) || !(remap[0] > 0.f) || !(remap[1] > 0.f) || !(remap[2] > 0.f))
    return fail(h, MDWT_ERR_ARG, "bounds must be finite and positive");
  h->remap = {remap[0], remap[1], remap[2]};
  return MDWT_OK;
}

int mdwt_init(mdwt_renderer* h, int mode, const int mc[3], int auto_camera) {
  if (!h) return MDWT_ERR_ARG;
  if (h->inited) return fail(h, MDWT_ERR_STATE, "already initialized");
  if (h->C == 0) return fail(h, MDWT_ERR_STATE, "no data source (push a channel first)");
  if (mode < 0 || mode > 9)
    return fail(h, MDWT_ERR_ARG,
                "mode must be 0-9 (0 N-DDA, 1 multi-majorant, 2 syncDDA, 3 cumulative, "
                "4/5 max/mix blend tracking, 6/7/8 ray marchers, 9 composite marcher)");
  int m[3];
  for (int i = 0; i < 3; ++i) {
    const int v = mc ? mc[i] : 0;
    if (v < 0) return fail(h, MDWT_ERR_ARG, "macrocell count must be positive");
    if (v > 0) {
      m[i] = v;
    } else {  // CalculateMCGridDims(4), renderer.cpp:1342-1355
      unsigned acc = 0;
      for (int c = 0; c < h->C; ++c) acc += unsigned(float(h->dims[i]) / 4.f);
      const unsigned avg = unsigned(float(acc) / float(h->C));
      m[i] = avg < 1 ? 1 : int(avg);
    }
  }
  const uint64_t n_cells = uint64_t(m[0]) * m[1] * m[2];
  if (n_cells * uint64_t(h->C) > (1ull << 28))
    return fail(h, MDWT_ERR_LIMIT, "macrocell grid too large");
  std::vector<int> wx, wy, wz;
  if (!axis_windows(h->dims[0], m[0], 0.f, h->remap.x, wx) ||
      !axis_windows(h->dims[1], m[1], 0.f, h->remap.y, wy) ||
      !axis_windows(h->dims[2], m[2], 0.f, h->remap.z, wz))
    return fail(h, MDWT_ERR_ARG, "macrocell projection windows are not intervals");

  HIPCHK(h, hipSetDevice(h->device));
  if (!h->hip_ready) {
    HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreate(&h->ev0));
    HIPCHK(h, hipEventCreate(&h->ev1));
    h->hip_ready = true;
  }
  for (int i = 0; i < 3; ++i) h->mc[i] = m[i];
  h->mode = mode;
  const size_t chan_bytes = size_t(h->dims[0]) * h->dims[1] * h->dims[2] * bytes_per_voxel(h->dtype);
  h->chan_stride = (chan_bytes + 255) & ~size_t(255);
  if (hipMalloc(&h->d_vol, h->chan_stride * h->C) != hipSuccess)
    return fail(h, MDWT_ERR_OOM, "device allocation of the volume failed");
  for (int c = 0; c < h->C; ++c)
    HIPCHK(h, hipMemcpyAsync(h->d_vol + size_t(c) * h->chan_stride, h->staging[c].data(), chan_bytes,
                             hipMemcpyHostToDevice, h->stream));
  int rc0 = build_range_grid(h, wx, wy, wz);
  if (rc0 != MDWT_OK) return rc0;
  if ((rc0 = make_bricks(h)) != MDWT_OK) return rc0;
  HIPCHK(h, hipMalloc(&h->d_counters, 4 * sizeof(unsigned long long)));
  HIPCHK(h, hipDeviceGetAttribute(&h->cu_count, hipDeviceAttributeMultiprocessorCount, h->device));
  // this device's copy of the allocation-free read's staging array
  HIPCHK(h, mdwt::fb_scratch_setup(h->device));
  // the reference drops its scalar buffers after the build (renderer.cpp:738-739)
  std::vector<std::vector<uint8_t>>().swap(h->staging);
  h->cm.assign(h->C, std::vector<Float4>());
  if (auto_camera) {  // renderer.cpp:856-863
    set_orientation(h, F3{0.5f, 0.5f, 0.5f + 2.5f}, F3{0.5f, 0.5f, 0.5f}, F3{0, 1, 0}, 45.0f);
  }
  h->inited = true;
  update_camera(h);
  int rc = alloc_frame(h);
  if (rc != MDWT_OK) return rc;
  h->frame_id = 0;
  reset_acc(h);
  return mdwt_reset_dt(h);
}

int mdwt_set_mode(mdwt_renderer* h, int mode) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (mode < 0 || mode > 9)
    return fail(h, MDWT_ERR_ARG,
                "mode must be 0-9 (0 N-DDA, 1 multi-majorant, 2 syncDDA, 3 cumulative, "
                "4/5 max/mix blend tracking, 6/7/8 ray marchers, 9 composite marcher)");
  HIPCHK(h, hipSetDevice(h->device));
  h->mode = mode;
  reset_acc(h);
  return rebuild_majorants(h);
}

int mdwt_set_xf_colormap(mdwt_renderer* h, int ch, const float* rgba, int n_texels) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (ch < 0 || ch >= h->C) return fail(h, MDWT_ERR_ARG, "channel id out of range");
  if (!rgba || n_texels < 1 || n_texels > 4096)
    return fail(h, MDWT_ERR_ARG, "colormap needs 1..4096 texels");
  std::vector<Float4> t(n_texels);
  for (int i = 0; i < n_texels; ++i) {
    const float* p = rgba + 4 * size_t(i);
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) || !std::isfinite(p[3]))
      return fail(h, MDWT_ERR_ARG, "non-finite colormap entry");
    // opacities come from a [0,1] widget in the reference; a negative one
    // would make free-flight steps run backwards
    if (p[3] < 0.f || p[3] > 1.f) return fail(h, MDWT_ERR_ARG, "colormap alpha must lie in [0, 1]");
    t[i] = Float4{p[0], p[1], p[2], std::pow(p[3], 3.f)};  // renderer.cpp:1098-1101
  }
  h->cm[ch] = std::move(t);
  h->cm_set[ch] = true;
  reset_acc(h);
  HIPCHK(h, hipSetDevice(h->device));
  return rebuild_majorants(h);
}

int mdwt_set_xf_opacity_scale(mdwt_renderer* h, int ch, float opacity_scale) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (ch < 0 || ch >= h->C) return fail(h, MDWT_ERR_ARG, "channel id out of range");
  if (!(opacity_scale >= 0.f && opacity_scale <= 1000.f))
    return fail(h, MDWT_ERR_ARG, "opacity scale must lie in [0, 1000]");
  h->os[ch] = opacity_scale;
  reset_acc(h);
  HIPCHK(h, hipSetDevice(h->device));
  return rebuild_majorants(h);
}

int mdwt_set_xf_range(mdwt_renderer* h, int ch, float xf_lo, float xf_hi) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (ch < 0 || ch >= h->C) return fail(h, MDWT_ERR_ARG, "channel id out of range");
  if (!std::isfinite(xf_lo) || !std::isfinite(xf_hi) || xf_lo == xf_hi)
    return fail(h, MDWT_ERR_ARG, "transfer-function range must be finite and non-empty");
  h->xlo[ch] = xf_lo;
  h->xhi[ch] = xf_hi;
  reset_acc(h);
  HIPCHK(h, hipSetDevice(h->device));
  return rebuild_majorants(h);
}

int mdwt_set_tf(mdwt_renderer* h, int ch, const float* rgba, int n_texels, float opacity_scale,
                float xf_lo, float xf_hi) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (ch < 0 || ch >= h->C) return fail(h, MDWT_ERR_ARG, "channel id out of range");
  if (!(opacity_scale >= 0.f && opacity_scale <= 1000.f))
    return fail(h, MDWT_ERR_ARG, "opacity scale must lie in [0, 1000]");
  if (!std::isfinite(xf_lo) || !std::isfinite(xf_hi) || xf_lo == xf_hi)
    return fail(h, MDWT_ERR_ARG, "transfer-function range must be finite and non-empty");
  h->os[ch] = opacity_scale;
  h->xlo[ch] = xf_lo;
  h->xhi[ch] = xf_hi;
  return mdwt_set_xf_colormap(h, ch, rgba, n_texels);
}

int mdwt_set_camera(mdwt_renderer* h, const float from[3], const float at[3], const float up[3],
                    float fovy_deg) {
  if (!h) return MDWT_ERR_ARG;
  if (!from || !at || !up || !finite3(from) || !finite3(at) || !finite3(up) ||
      !std::isfinite(fovy_deg))
    return fail(h, MDWT_ERR_ARG, "camera vectors must be finite");
  set_orientation(h, F3{from[0], from[1], from[2]}, F3{at[0], at[1], at[2]}, F3{up[0], up[1], up[2]},
                  fovy_deg);
  if (h->inited) update_camera(h);
  return MDWT_OK;
}

int mdwt_set_dt(mdwt_renderer* h, float dt) {
  if (!h) return MDWT_ERR_ARG;
  if (!std::isfinite(dt)) return fail(h, MDWT_ERR_ARG, "non-finite dt");
  h->dt = std::fmax(dt, 1e-4f);  // renderer.cpp:1460-1466
  reset_acc(h);
  return MDWT_OK;
}

int mdwt_reset_dt(mdwt_renderer* h) {
  if (!h) return MDWT_ERR_ARG;
  if (h->C == 0) return fail(h, MDWT_ERR_STATE, "no data source");
  // renderer.cpp:1444-1453: the loop assigns, so the last channel wins
  float minVoxelSideLength = 3.402823466e38f;
  for (int c = 0; c < h->C; ++c)
    minVoxelSideLength =
        std::fmin(h->remap.x / float(h->dims[0]),
                  std::fmin(h->remap.y / float(h->dims[1]), h->remap.z / float(h->dims[2])));
  return mdwt_set_dt(h, minVoxelSideLength * 0.5f);
}

int mdwt_set_frame(mdwt_renderer* h, int width, int height, int spp, int accumulate,
                   const float bg[3]) {
  if (!h) return MDWT_ERR_ARG;
  if (width <= 0 || height <= 0 || spp <= 0) return fail(h, MDWT_ERR_ARG, "frame sizes must be positive");
  // pixel indices are 32-bit in the kernels: at most 2^30 pixels per frame
  if (width > 16384 * 4 || height > 16384 * 4 || uint64_t(width) * uint64_t(height) > (1ull << 30))
    return fail(h, MDWT_ERR_LIMIT, "frame too large (at most 2^30 pixels, 65536 per side)");
  if (bg && !finite3(bg)) return fail(h, MDWT_ERR_ARG, "non-finite background");
  const bool resize = width != h->W || height != h->H;
  h->W = width;
  h->H = height;
  h->spp = spp;
  h->accumulate = accumulate != 0;
  if (bg)
    for (int i = 0; i < 3; ++i) h->bg[i] = bg[i];
  reset_acc(h);
  if (h->inited) {
    update_camera(h);
    if (resize) {
      HIPCHK(h, hipSetDevice(h->device));
      return alloc_frame(h);
    }
  }
  return MDWT_OK;
}

int mdwt_set_lighting(mdwt_renderer* h, float light_intensity, float ambient) {
  if (!h) return MDWT_ERR_ARG;
  if (!std::isfinite(light_intensity) || !std::isfinite(ambient))
    return fail(h, MDWT_ERR_ARG, "non-finite lighting");
  h->light = light_intensity;
  h->ambient = ambient;
  reset_acc(h);
  return MDWT_OK;
}

int mdwt_set_light(mdwt_renderer* h, const float light_dir[3], int shadows, int gradient_shading) {
  if (!h || !light_dir) return MDWT_ERR_ARG;
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(light_dir[i])) return fail(h, MDWT_ERR_ARG, "non-finite light direction");
  for (int i = 0; i < 3; ++i) h->sh.light_dir[i] = light_dir[i];
  h->sh.shadows = shadows ? 1 : 0;
  h->sh.gradient = gradient_shading ? 1 : 0;
  reset_acc(h);
  return MDWT_OK;
}

int mdwt_set_material(mdwt_renderer* h, int channel, float diffuse, float specular,
                      float shininess, float gradient_step, const float specular_color[3]) {
  if (!h || !specular_color) return MDWT_ERR_ARG;
  if (channel < 0 || channel >= mdwt::kMaxChannels)
    return fail(h, MDWT_ERR_ARG, "channel out of range");
  const float v[7] = {diffuse, specular, shininess, gradient_step, specular_color[0],
                      specular_color[1], specular_color[2]};
  for (float x : v)
    if (!std::isfinite(x)) return fail(h, MDWT_ERR_ARG, "non-finite material value");
  if (!(gradient_step > 0.f)) return fail(h, MDWT_ERR_ARG, "gradient step must be positive");
  h->sh.diffuse[channel] = diffuse;
  h->sh.specular[channel] = specular;
  h->sh.shininess[channel] = shininess;
  h->sh.grad_step[channel] = gradient_step;
  for (int i = 0; i < 3; ++i) h->sh.spec[channel][i] = specular_color[i];
  reset_acc(h);
  return MDWT_OK;
}

int mdwt_material_preset(const char* name, float out[7]) {
  if (!name || !out) return MDWT_ERR_ARG;
  // renderer.cpp:1250-1301 ApplyMaterialPreset (unknown names: the default)
  struct Preset {
    const char* name;
    float v[7];
  };
  static const Preset kPresets[] = {
      {"Plastic", {0.1f, 0.50f, 32.0f, 0.01f, 0.5f, 0.5f, 0.5f}},
      {"Iron", {0.51f, 0.51f, 51.0f, 0.01f, 0.51f, 0.51f, 0.51f}},
      {"Chrome", {0.4f, 0.77f, 77.0f, 0.01f, 0.77f, 0.77f, 0.77f}},
      {"Emerald", {0.255f, 0.66f, 77.0f, 0.005f, 0.63f, 0.73f, 0.63f}},
  };
  static const float kDefault[7] = {0.8f, 0.3f, 8.0f, 0.01f, 1.0f, 1.0f, 1.0f};
  const float* v = kDefault;
  for (const Preset& p : kPresets)
    if (std::strcmp(p.name, name) == 0) v = p.v;
  for (int i = 0; i < 7; ++i) out[i] = v[i];
  return MDWT_OK;
}

int mdwt_set_seed(mdwt_renderer* h, uint32_t seed) {
  if (!h) return MDWT_ERR_ARG;
  h->seed = seed;
  reset_acc(h);
  return MDWT_OK;
}

int mdwt_render(mdwt_renderer* h, int n_frames, float* ms_out) {
  if (!h) return MDWT_ERR_ARG;
  int rc = ready_to_render(h);
  if (rc != MDWT_OK) return rc;
  if (n_frames < 0) return fail(h, MDWT_ERR_ARG, "negative frame count");
  HIPCHK(h, hipSetDevice(h->device));
  if (h->order_mode == 2 && h->need_resort && h->cost_ready) {
    rc = resort_blocks(h);
    if (rc != MDWT_OK) return rc;
    h->need_resort = false;
  }
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int f = 0; f < n_frames; ++f) {
    rc = launch_frame(h, false);
    if (rc != MDWT_OK) return rc;
    h->accum_id++;
    h->frame_id++;
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_kernel_ms = n_frames > 0 ? ms / float(n_frames) : 0.f;
  if (ms_out) *ms_out = ms;
  return MDWT_OK;
}

int mdwt_reset_accumulation(mdwt_renderer* h) {
  if (!h) return MDWT_ERR_ARG;
  reset_acc(h);
  return MDWT_OK;
}

int mdwt_read_accum(mdwt_renderer* h, float* rgba, float* m2) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t bytes = h->fb_pixels * sizeof(Float4);
  if (m2 && !h->d_m2) return fail(h, MDWT_ERR_STATE, "second moments are off (mdwt_set_moments)");
  if (rgba) HIPCHK(h, hipMemcpyAsync(rgba, h->d_accum, bytes, hipMemcpyDeviceToHost, h->stream));
  if (m2) HIPCHK(h, hipMemcpyAsync(m2, h->d_m2, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return MDWT_OK;
}

int mdwt_read_fb(mdwt_renderer* h, uint32_t* rgba8) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (!rgba8) return fail(h, MDWT_ERR_ARG, "null output");
  HIPCHK(h, hipSetDevice(h->device));
  uint32_t* d = nullptr;
  HIPCHK(h, hipMalloc(&d, h->fb_pixels * sizeof(uint32_t)));
  hipError_t e = mdwt::launch_rgba8(h->d_accum, d, int(h->fb_pixels), h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rgba8, d, h->fb_pixels * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(h, MDWT_ERR_HIP, std::string("read_fb: ") + hipGetErrorString(e));
  return MDWT_OK;
}

int mdwt_read_fb_noalloc(mdwt_renderer* h, uint32_t* rgba8) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (!rgba8) return fail(h, MDWT_ERR_ARG, "null output");
  if (h->fb_pixels > mdwt::kFbScratchPixels)
    return fail(h, MDWT_ERR_LIMIT,
                "read_fb_noalloc: frame of " + std::to_string(h->fb_pixels) +
                    " pixels exceeds the static staging area of " +
                    std::to_string(mdwt::kFbScratchPixels) + " pixels");
  HIPCHK(h, hipSetDevice(h->device));
  uint32_t* d = mdwt::fb_scratch(h->device);  // set up by mdwt_init on this device
  if (!d) return fail(h, MDWT_ERR_STATE, "read_fb_noalloc: staging area not set up on this device");
  hipError_t e = mdwt::launch_rgba8(h->d_accum, d, int(h->fb_pixels), h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rgba8, d, h->fb_pixels * sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess)
    return fail(h, MDWT_ERR_HIP, std::string("read_fb_noalloc: ") + hipGetErrorString(e));
  return MDWT_OK;
}

int mdwt_read_ranges(mdwt_renderer* h, float* out) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n = size_t(h->mc[0]) * h->mc[1] * h->mc[2] * h->C * 2;
  HIPCHK(h, hipMemcpy(out, h->d_ranges, n * sizeof(float), hipMemcpyDeviceToHost));
  return MDWT_OK;
}

int mdwt_read_majorants(mdwt_renderer* h, float* out) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (!h->maj_valid) return fail(h, MDWT_ERR_STATE, "majorants not built (set every transfer function)");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n = size_t(h->mc[0]) * h->mc[1] * h->mc[2] * (h->mode <= 2 ? h->C : 1);
  HIPCHK(h, hipMemcpy(out, h->d_maj, n * sizeof(float), hipMemcpyDeviceToHost));
  return MDWT_OK;
}

int mdwt_get_info(mdwt_renderer* h, int out[10]) {
  if (!h || !out) return MDWT_ERR_ARG;
  out[0] = h->mc[0];
  out[1] = h->mc[1];
  out[2] = h->mc[2];
  out[3] = h->dims[0];
  out[4] = h->dims[1];
  out[5] = h->dims[2];
  out[6] = h->C;
  out[7] = h->mode;
  out[8] = int(h->frame_id);
  out[9] = int(h->accum_id);
  return MDWT_OK;
}

int mdwt_get_dt(mdwt_renderer* h, float* dt) {
  if (!h || !dt) return MDWT_ERR_ARG;
  *dt = h->dt;
  return MDWT_OK;
}

int mdwt_get_camera(mdwt_renderer* h, float out[12]) {
  if (!h || !out) return MDWT_ERR_ARG;
  const float* src[4] = {h->cam.org, h->cam.llc, h->cam.hor, h->cam.ver};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) out[3 * i + k] = src[i][k];
  return MDWT_OK;
}

int mdwt_get_domains(mdwt_renderer* h, float* out) {
  if (!h || !out) return MDWT_ERR_ARG;
  for (int c = 0; c < h->C; ++c) {
    out[2 * c] = h->vmin[c];
    out[2 * c + 1] = h->vmax[c];
  }
  return MDWT_OK;
}

int mdwt_trace_cells(mdwt_renderer* h, const float* org, const float* dir, int n, int max_cells,
                     int32_t* cells, float* t01, int32_t* counts) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (n < 0 || max_cells < 0 || (n > 0 && (!org || !dir || !counts)) ||
      (n > 0 && max_cells > 0 && (!cells || !t01)))
    return fail(h, MDWT_ERR_ARG, "bad trace_cells arguments");
  if (n == 0) return MDWT_OK;
  HIPCHK(h, hipSetDevice(h->device));
  const size_t nc = size_t(n) * std::max(max_cells, 1);
  float *d_org = nullptr, *d_dir = nullptr, *d_t = nullptr;
  int *d_cells = nullptr, *d_cnt = nullptr;
  hipError_t e = hipMalloc(&d_org, size_t(n) * 12);
  if (e == hipSuccess) e = hipMalloc(&d_dir, size_t(n) * 12);
  if (e == hipSuccess) e = hipMalloc(&d_t, nc * 8);
  if (e == hipSuccess) e = hipMalloc(&d_cells, nc * 4);
  if (e == hipSuccess) e = hipMalloc(&d_cnt, size_t(n) * 4);
  if (e == hipSuccess) e = hipMemcpy(d_org, org, size_t(n) * 12, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_dir, dir, size_t(n) * 12, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const mdwt::VolumeParams v = volume_params(h);
    e = mdwt::launch_trace_cells(v, h->mc, d_org, d_dir, n, max_cells, d_cells, d_t, d_cnt, h->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && max_cells > 0) {
    e = hipMemcpy(cells, d_cells, size_t(n) * max_cells * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(t01, d_t, size_t(n) * max_cells * 8, hipMemcpyDeviceToHost);
  }
  if (e == hipSuccess) e = hipMemcpy(counts, d_cnt, size_t(n) * 4, hipMemcpyDeviceToHost);
  void* ptrs[] = {d_org, d_dir, d_t, d_cells, d_cnt};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return fail(h, MDWT_ERR_HIP, std::string("trace_cells: ") + hipGetErrorString(e));
  return MDWT_OK;
}

int mdwt_gen_rays(mdwt_renderer* h, uint32_t frame, uint32_t sample, float* org, float* dir) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (!org || !dir) return fail(h, MDWT_ERR_ARG, "null output");
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n = size_t(h->W) * h->H;
  float *d_org = nullptr, *d_dir = nullptr;
  hipError_t e = hipMalloc(&d_org, n * 12);
  if (e == hipSuccess) e = hipMalloc(&d_dir, n * 12);
  if (e == hipSuccess)
    e = mdwt::launch_gen_rays(h->cam, h->W, h->H, h->seed, frame, sample, h->jitter2d ? 1 : 0,
                              d_org, d_dir, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(org, d_org, n * 12, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(dir, d_dir, n * 12, hipMemcpyDeviceToHost);
  if (d_org) (void)hipFree(d_org);
  if (d_dir) (void)hipFree(d_dir);
  if (e != hipSuccess) return fail(h, MDWT_ERR_HIP, std::string("gen_rays: ") + hipGetErrorString(e));
  return MDWT_OK;
}

int mdwt_clip_rays(mdwt_renderer* h, const float* org, const float* dir, int n, float* t01,
                   int32_t* hit) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (n < 0 || (n > 0 && (!org || !dir || !t01 || !hit)))
    return fail(h, MDWT_ERR_ARG, "bad clip_rays arguments");
  if (n == 0) return MDWT_OK;
  HIPCHK(h, hipSetDevice(h->device));
  float *d_org = nullptr, *d_dir = nullptr, *d_t = nullptr;
  int* d_hit = nullptr;
  hipError_t e = hipMalloc(&d_org, size_t(n) * 12);
  if (e == hipSuccess) e = hipMalloc(&d_dir, size_t(n) * 12);
  if (e == hipSuccess) e = hipMalloc(&d_t, size_t(n) * 8);
  if (e == hipSuccess) e = hipMalloc(&d_hit, size_t(n) * 4);
  if (e == hipSuccess) e = hipMemcpy(d_org, org, size_t(n) * 12, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_dir, dir, size_t(n) * 12, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = mdwt::launch_clip(volume_params(h), d_org, d_dir, n, d_t, d_hit, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(t01, d_t, size_t(n) * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hit, d_hit, size_t(n) * 4, hipMemcpyDeviceToHost);
  void* ptrs[] = {d_org, d_dir, d_t, d_hit};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return fail(h, MDWT_ERR_HIP, std::string("clip_rays: ") + hipGetErrorString(e));
  return MDWT_OK;
}

int mdwt_count_frame(mdwt_renderer* h, uint64_t out[4]) {
  if (!h || !out) return MDWT_ERR_ARG;
  int rc = ready_to_render(h);
  if (rc != MDWT_OK) return rc;
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipMemsetAsync(h->d_counters, 0, 4 * sizeof(unsigned long long), h->stream));
  rc = launch_frame(h, true);
  if (rc != MDWT_OK) return rc;
  unsigned long long c[4];
  HIPCHK(h, hipMemcpyAsync(c, h->d_counters, sizeof(c), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int i = 0; i < 4; ++i) out[i] = c[i];
  return MDWT_OK;
}

int mdwt_last_kernel_ms(mdwt_renderer* h, float* ms) {
  if (!h || !ms) return MDWT_ERR_ARG;
  *ms = h->last_kernel_ms;
  return MDWT_OK;
}

int mdwt_set_tile_order(mdwt_renderer* h, int order, int tile) {
  if (!h) return MDWT_ERR_ARG;
  if (order < 0 || order > 2)
    return fail(h, MDWT_ERR_ARG, "order must be 0 (rows), 1 (Hilbert) or 2 (cost-sorted)");
  if (tile != 8 && tile != 16 && tile != 32) return fail(h, MDWT_ERR_ARG, "tile must be 8, 16 or 32");
  h->tile = tile;
  h->order_mode = order;
  return reorder(h);
}

int mdwt_set_schedule(mdwt_renderer* h, int schedule) {
  if (!h) return MDWT_ERR_ARG;
  if (schedule < 0 || schedule > 1)
    return fail(h, MDWT_ERR_ARG, "schedule must be 0 (thread per pixel) or 1 (wave-coherent)");
  h->schedule = schedule;
  return MDWT_OK;
}

int mdwt_set_conservative_majorants(mdwt_renderer* h, int on) {
  if (!h) return MDWT_ERR_ARG;
  if (h->conservative_maj == (on != 0)) return MDWT_OK;
  h->conservative_maj = on != 0;
  if (!h->inited) return MDWT_OK;
  HIPCHK(h, hipSetDevice(h->device));
  reset_acc(h);
  return rebuild_majorants(h);
}

int mdwt_set_jitter_2d(mdwt_renderer* h, int on) {
  if (!h) return MDWT_ERR_ARG;
  h->jitter2d = on != 0;
  return MDWT_OK;
}

int mdwt_set_moments(mdwt_renderer* h, int on) {
  if (!h) return MDWT_ERR_ARG;
  const bool want = on != 0;
  if (want == h->moments) return MDWT_OK;
  h->moments = want;
  reset_acc(h);
  if (!h->inited) return MDWT_OK;
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (h->fb_pixels) {
    const int rc = alloc_frame(h);
    if (rc != MDWT_OK) return rc;
  }
  if (!h->views.empty()) {
    for (auto& v : h->views) {
      const size_t bytes = size_t(v.width) * v.height * sizeof(Float4);
      if (want && !v.m2) {
        if (hipMalloc(&v.m2, bytes) != hipSuccess) {
          v.m2 = nullptr;
          return fail(h, MDWT_ERR_OOM, "view second moments");
        }
        HIPCHK(h, hipMemset(v.m2, 0, bytes));
      } else if (!want && v.m2) {
        (void)hipFree(v.m2);
        v.m2 = nullptr;
      }
    }
    HIPCHK(h, hipMemcpy(h->d_views, h->views.data(), h->views.size() * sizeof(mdwt::ViewParams),
                        hipMemcpyHostToDevice));
    h->batch_accum = 0;
  }
  return MDWT_OK;
}

int mdwt_stream_read(mdwt_renderer* h, uint64_t bytes, int iters, float* gbps) {
  if (!h || !gbps) return MDWT_ERR_ARG;
  if (bytes < 16 || iters < 1) return fail(h, MDWT_ERR_ARG, "bad read size");
  if (!h->hip_ready) {
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreate(&h->ev0));
    HIPCHK(h, hipEventCreate(&h->ev1));
    h->hip_ready = true;
  }
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n16 = size_t(bytes / 16);
  const unsigned grid = mdwt::kReadGrid;
  void *a = nullptr, *sums = nullptr;
  if (hipMalloc(&a, n16 * 16) != hipSuccess) return fail(h, MDWT_ERR_OOM, "read buffer");
  if (hipMalloc(&sums, size_t(grid) * sizeof(float)) != hipSuccess) {
    (void)hipFree(a);
    return fail(h, MDWT_ERR_OOM, "read partial sums");
  }
  hipError_t e = hipMemsetAsync(a, 1, n16 * 16, h->stream);
  if (e == hipSuccess) e = mdwt::launch_read(a, n16, static_cast<float*>(sums), h->stream);
  if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
  for (int i = 0; i < iters && e == hipSuccess; ++i)
    e = mdwt::launch_read(a, n16, static_cast<float*>(sums), h->stream);
  if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
  (void)hipFree(a);
  (void)hipFree(sums);
  if (e != hipSuccess) return fail(h, MDWT_ERR_HIP, std::string("stream_read: ") + hipGetErrorString(e));
  *gbps = float(double(n16 * 16) * iters / (double(ms) * 1e-3) / 1e9);
  return MDWT_OK;
}

int mdwt_stream_copy(mdwt_renderer* h, uint64_t bytes, int iters, float* gbps) {
  if (!h || !gbps) return MDWT_ERR_ARG;
  if (bytes < 16 || iters < 1) return fail(h, MDWT_ERR_ARG, "bad copy size");
  if (!h->hip_ready) {
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreate(&h->ev0));
    HIPCHK(h, hipEventCreate(&h->ev1));
    h->hip_ready = true;
  }
  HIPCHK(h, hipSetDevice(h->device));
  const size_t n16 = size_t(bytes / 16);
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n16 * 16) != hipSuccess) return fail(h, MDWT_ERR_OOM, "copy buffer");
  if (hipMalloc(&b, n16 * 16) != hipSuccess) {
    (void)hipFree(a);
    return fail(h, MDWT_ERR_OOM, "copy buffer");
  }
  hipError_t e = hipMemsetAsync(a, 1, n16 * 16, h->stream);
  if (e == hipSuccess) e = mdwt::launch_copy(a, b, n16, h->stream);
  if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
  for (int i = 0; i < iters && e == hipSuccess; ++i) e = mdwt::launch_copy(a, b, n16, h->stream);
  if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
  if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return fail(h, MDWT_ERR_HIP, std::string("stream_copy: ") + hipGetErrorString(e));
  *gbps = float(2.0 * double(n16 * 16) * iters / (double(ms) * 1e-3) / 1e9);
  return MDWT_OK;
}

int mdwt_hilbert_d2xy(int bits, uint32_t d, uint32_t* x, uint32_t* y) {
  if (bits < 0 || bits > 15 || !x || !y) return MDWT_ERR_ARG;
  if (bits < 16 && uint64_t(d) >= (uint64_t(1) << (2 * bits))) return MDWT_ERR_ARG;
  hilbert_d2xy(bits, d, *x, *y);
  return MDWT_OK;
}

}  // extern "C"

// ===================================================== batched views ======
namespace {

void free_batch(mdwt_renderer* h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  h->graph_valid = false;
  for (auto& v : h->views) {
    if (v.accum) (void)hipFree(v.accum);
    if (v.m2) (void)hipFree(v.m2);
  }
  h->views.clear();
  void* ptrs[] = {h->d_views, h->d_items, h->d_pitems, h->d_work, h->d_item_ticks};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  h->d_views = nullptr;
  h->d_items = nullptr;
  h->d_pitems = nullptr;
  h->pitems.clear();
  h->pitem_ticks.clear();
  h->pitems_measured = false;
  h->d_work = nullptr;
  h->d_item_ticks = nullptr;
  h->items.clear();
  for (hipStream_t q : h->batch_streams) {
    (void)hipStreamSynchronize(q);
    (void)hipStreamDestroy(q);
  }
  h->batch_streams.clear();
  for (uint32_t* q : h->view_order)
    if (q) (void)hipFree(q);
  for (auto* v : {&h->view_ev, &h->view_ev_end, &h->join_ev}) {
    for (hipEvent_t e : *v)
      if (e) (void)hipEventDestroy(e);
    v->clear();
  }
  h->view_order.clear();
  h->view_blocks.clear();
  h->view_us.clear();
}

// Per-view block orders (Hilbert tiles of the current block shape) for the
// per-view launches.
int build_view_orders(mdwt_renderer* h) {
  for (uint32_t* q : h->view_order)
    if (q) (void)hipFree(q);
  h->view_order.assign(h->views.size(), nullptr);
  h->view_blocks.assign(h->views.size(), 0);
  for (size_t i = 0; i < h->views.size(); ++i) {
    const auto& vw = h->views[i];
    const int nbx = (vw.width + h->bw - 1) / h->bw, nby = (vw.height + h->bh - 1) / h->bh;
    const std::vector<uint32_t> order = block_order(nbx, nby, h->bw, h->bh, h->tile, true);
    HIPCHK(h, hipMalloc(&h->view_order[i], order.size() * sizeof(uint32_t)));
    HIPCHK(h, hipMemcpy(h->view_order[i], order.data(), order.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice));
    h->view_blocks[i] = int(order.size());
  }
  return MDWT_OK;
}

// Work items of all views: views by decreasing work (pixels x spp), each
// view's blocks in the Hilbert-tile order of the current block shape.
int build_items(mdwt_renderer* h) {
  std::vector<int> idx(h->views.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = int(i);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    const auto& va = h->views[size_t(a)];
    const auto& vb = h->views[size_t(b)];
    return uint64_t(va.width) * va.height * va.spp > uint64_t(vb.width) * vb.height * vb.spp;
  });
  std::vector<uint32_t> items;
  for (int v : idx) {
    const auto& vw = h->views[size_t(v)];
    const int nbx = (vw.width + h->bw - 1) / h->bw, nby = (vw.height + h->bh - 1) / h->bh;
    if (nbx > 1024 || nby > 1024) return fail(h, MDWT_ERR_LIMIT, "view has more than 1024 blocks per side");
    for (uint32_t o : block_order(nbx, nby, h->bw, h->bh, h->tile, true))
      items.push_back((uint32_t(v) << 20) | ((o >> 16) << 10) | (o & 0xffffu));
  }
  if (h->d_items) (void)hipFree(h->d_items);
  if (h->d_pitems) (void)hipFree(h->d_pitems);
  if (h->d_item_ticks) (void)hipFree(h->d_item_ticks);
  h->d_items = nullptr;
  h->d_pitems = nullptr;
  h->d_item_ticks = nullptr;
  HIPCHK(h, hipMalloc(&h->d_items, items.size() * sizeof(uint32_t)));
  HIPCHK(h, hipMalloc(&h->d_pitems, items.size() * sizeof(uint32_t)));
  HIPCHK(h, hipMalloc(&h->d_item_ticks, 2 * items.size() * sizeof(unsigned long long)));
  HIPCHK(h, hipMemcpy(h->d_items, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->d_pitems, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(h->d_item_ticks, 0, 2 * items.size() * sizeof(unsigned long long)));
  h->pitems = items;
  h->pitem_ticks.clear();
  h->pitems_measured = false;
  h->items = std::move(items);
  return build_view_orders(h);
}

// The single-view launch of view i at the batch's frame and accumulation ids.
mdwt::RenderParams view_params(const mdwt_renderer* h, size_t i) {
  mdwt::RenderParams p = render_params(h);
  const mdwt::ViewParams& v = h->views[i];
  p.cam = v.cam;
  p.fr.width = v.width;
  p.fr.height = v.height;
  p.fr.spp = v.spp;
  p.fr.seed = v.seed;
  p.fr.frame_id = h->batch_frame;
  p.fr.accum_id = h->batch_accum;
  p.fr.accum = v.accum;
  p.fr.m2 = v.m2;
  p.fr.block_order = h->view_order[i];
  p.fr.n_blocks = h->view_blocks[i];
  p.fr.schedule = eff_schedule(h);
  p.fr.wave_cost = nullptr;
  p.fr.aov = nullptr;  // batched views have no auxiliary buffers
  return p;
}

// Streams the views of a batch are spread over (round-robin, default
// priority): small views do not fill the GPU one launch at a time.
constexpr int kBatchStreams = 4;

int make_events(mdwt_renderer* h, std::vector<hipEvent_t>& v, size_t n) {
  if (v.size() == n) return MDWT_OK;
  for (hipEvent_t e : v)
    if (e) (void)hipEventDestroy(e);
  v.assign(n, nullptr);
  for (auto& e : v) HIPCHK(h, hipEventCreate(&e));
  return MDWT_OK;
}

// n_frames frames of every view, one launch per view per frame, view i on
// batch stream i % kBatchStreams (a view's frames stay in order on its
// stream). The batch streams start after the handle's stream has reached
// this call and the handle's stream waits for all of them at the end, so
// the call is ordered like any other on the handle. Events around the last
// frame's launches give the per-view times.
int batch_render_views(mdwt_renderer* h, int n_frames, float* ms_out) {
  const size_t n = h->views.size();
  int rc;
  if ((rc = make_events(h, h->view_ev, n)) != MDWT_OK ||
      (rc = make_events(h, h->view_ev_end, n)) != MDWT_OK ||
      (rc = make_events(h, h->join_ev, kBatchStreams)) != MDWT_OK)
    return rc;
  while (h->batch_streams.size() < size_t(kBatchStreams)) {
    hipStream_t q = nullptr;
    HIPCHK(h, hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
    h->batch_streams.push_back(q);
  }
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (hipStream_t q : h->batch_streams) HIPCHK(h, hipStreamWaitEvent(q, h->ev0, 0));
  for (int f = 0; f < n_frames; ++f) {
    for (size_t i = 0; i < n; ++i) {
      hipStream_t q = h->batch_streams[i % kBatchStreams];
      HIPCHK(h, hipEventRecord(h->view_ev[i], q));
      HIPCHK(h, mdwt::launch_render(view_params(h, i), h->mode, h->dtype, false, q));
      HIPCHK(h, hipEventRecord(h->view_ev_end[i], q));
    }
    h->batch_frame++;
    h->batch_accum = h->accumulate ? h->batch_accum + 1u : 0u;
  }
  for (int k = 0; k < kBatchStreams; ++k) {
    HIPCHK(h, hipEventRecord(h->join_ev[size_t(k)], h->batch_streams[size_t(k)]));
    HIPCHK(h, hipStreamWaitEvent(h->stream, h->join_ev[size_t(k)], 0));
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (n_frames > 0) {
    h->view_us.assign(n, 0.f);
    for (size_t i = 0; i < n; ++i) {
      float t = 0.f;
      HIPCHK(h, hipEventElapsedTime(&t, h->view_ev[i], h->view_ev_end[i]));
      h->view_us[i] = t * 1000.f;
    }
  }
  h->last_kernel_ms = n_frames > 0 ? ms / float(n_frames) : 0.f;
  if (ms_out) *ms_out = ms;
  return MDWT_OK;
}

// The one-launch form: every view's pixel blocks in one grid per frame, at
// the batch's frame and accumulation ids.
mdwt::RenderParams views_params(const mdwt_renderer* h) {
  mdwt::RenderParams p = render_params(h);
  p.fr.schedule = eff_schedule(h);
  p.fr.wave_cost = nullptr;
  p.fr.views = h->d_views;
  p.fr.items = h->d_items;
  p.fr.n_items = int(h->items.size());
  p.fr.work = nullptr;
  p.fr.item_ticks = nullptr;
  p.fr.frame_id = h->batch_frame;
  p.fr.accum_id = h->batch_accum;
  p.fr.aov = nullptr;  // batched views have no auxiliary buffers
  return p;
}

// Per-view time of the last one-launch frame from its workgroups' wall-clock
// stamps (100 MHz; stamp i belongs to list[i]): from the first workgroup's
// start to the view's last workgroup's end. The stamps read stay in *keep
// when given.
int views_times_from_ticks(mdwt_renderer* h, const std::vector<uint32_t>& list,
                           std::vector<unsigned long long>* keep = nullptr) {
  const size_t n = list.size();
  std::vector<unsigned long long> t(2 * n);
  HIPCHK(h, hipMemcpy(t.data(), h->d_item_ticks, t.size() * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost));
  unsigned long long t0 = ~0ull;
  for (size_t i = 0; i < n; ++i) t0 = std::min(t0, t[2 * i]);
  std::vector<unsigned long long> end(h->views.size(), 0ull);
  for (size_t i = 0; i < n; ++i) {
    const size_t v = list[i] >> 20;
    end[v] = std::max(end[v], t[2 * i + 1]);
  }
  h->view_us.assign(h->views.size(), 0.f);
  for (size_t v = 0; v < end.size(); ++v)
    h->view_us[v] = end[v] >= t0 ? float(double(end[v] - t0) / 100.0) : 0.f;
  if (keep) keep->swap(t);
  return MDWT_OK;
}

// n_frames frames of every view, one launch per frame on the handle's stream;
// the last frame records its workgroups' wall clocks for the per-view times.
int batch_render_one_launch(mdwt_renderer* h, int n_frames, float* ms_out) {
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int f = 0; f < n_frames; ++f) {
    mdwt::RenderParams p = views_params(h);
    if (f == n_frames - 1) p.fr.item_ticks = h->d_item_ticks;
    HIPCHK(h, mdwt::launch_views(p, h->mode, h->dtype, false, h->stream));
    h->batch_frame++;
    h->batch_accum = h->accumulate ? h->batch_accum + 1u : 0u;
  }
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  if (n_frames > 0) {
    const int rc = views_times_from_ticks(h, h->items);
    if (rc != MDWT_OK) return rc;
  }
  h->last_kernel_ms = n_frames > 0 ? ms / float(n_frames) : 0.f;
  if (ms_out) *ms_out = ms;
  return MDWT_OK;
}

// Launch form 2: the one-launch parameters with the device frame and
// accumulation ids (the captured graph advances them itself) and the per-item
// wall clocks.
mdwt::RenderParams persistent_params(const mdwt_renderer* h) {
  mdwt::RenderParams p = views_params(h);
  p.fr.frame_id = 0;  // read from work[1] / work[2] instead
  p.fr.accum_id = 0;
  p.fr.items = h->d_pitems;
  p.fr.work = h->d_work;
  p.fr.item_ticks = h->d_item_ticks;
  return p;
}

// Re-deal form 2's items by the costs the last form-2 frame measured (its
// per-item wall-clock stamps, kept on the host): longest processing time
// first over the G resident workgroups (mdwt::deal_sources). The list
// changes in place: the captured graph reads it at each replay.
int deal_by_cost(mdwt_renderer* h, int G) {
  const size_t n = h->pitems.size();
  if (n == 0 || G <= 0 || h->pitem_ticks.size() != 2 * n) return MDWT_OK;
  const std::vector<uint32_t> src = mdwt::deal_sources(h->pitem_ticks.data(), n, size_t(G));
  std::vector<uint32_t> dealt(n);
  for (size_t idx = 0; idx < n; ++idx) dealt[idx] = h->pitems[src[idx]];
  h->pitems = std::move(dealt);
  HIPCHK(h, hipMemcpyAsync(h->d_pitems, h->pitems.data(), n * sizeof(uint32_t),
                           hipMemcpyHostToDevice, h->stream));
  return MDWT_OK;
}

// resident workgroups of the persistent launch (cached per kernel choice),
// at most one per item
int persistent_grid(mdwt_renderer* h, const mdwt::RenderParams& p, int* grid) {
  const bool lit = p.fr.sh.shadows || p.fr.sh.gradient || p.fr.aov != nullptr;
  const int key = int(mdwt::render_lds_bytes(p)) * 64 + h->mode * 16 + h->bw * h->bh / 64 * 2 +
                  (lit ? 1 << 30 : 0);
  if (key != h->batch_grid_key) {
    int per_cu = 0;
    HIPCHK(h, mdwt::views_persistent_blocks_per_cu(p, h->mode, h->dtype, &per_cu));
    h->batch_grid = std::max(1, per_cu) * h->cu_count;
    h->batch_grid_key = key;
  }
  *grid = std::min(h->batch_grid, std::max(1, p.fr.n_items));
  return MDWT_OK;
}

// n_frames frames of every view by graph replay: the persistent launch and
// the id advance are captured once (again whenever the parameters change)
// and the graph is replayed once per frame. The device ids are set from the
// host's before the replays, so frames split over calls and resets carry.
int batch_render_persistent(mdwt_renderer* h, int n_frames, float* ms_out) {
  const mdwt::RenderParams p = persistent_params(h);
  int grid = 0;
  int rc = persistent_grid(h, p, &grid);
  if (rc != MDWT_OK) return rc;
  if (!h->graph_valid || grid != h->graph_grid ||
      std::memcmp(&p, &h->graph_params, sizeof(p)) != 0) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_valid = false;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const hipError_t le = mdwt::launch_views_persistent(p, h->mode, h->dtype, grid, h->stream);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(h->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(h, MDWT_ERR_HIP, std::string("graph capture: ") +
                                       hipGetErrorString(le != hipSuccess ? le : ce));
    }
    h->graph = g;
    HIPCHK(h, hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
    h->graph_params = p;
    h->graph_grid = grid;
    h->graph_valid = true;
  }
  if (h->pitems_measured && (rc = deal_by_cost(h, grid)) != MDWT_OK) return rc;
  const unsigned ids[3] = {0u, h->batch_frame, h->batch_accum};
  HIPCHK(h, hipMemcpyAsync(h->d_work, ids, sizeof(ids), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  for (int f = 0; f < n_frames; ++f) HIPCHK(h, hipGraphLaunch(h->graph_exec, h->stream));
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->batch_frame += uint32_t(n_frames);
  h->batch_accum = h->accumulate ? h->batch_accum + uint32_t(n_frames) : 0u;
  if (n_frames > 0) {
    rc = views_times_from_ticks(h, h->pitems, &h->pitem_ticks);
    if (rc != MDWT_OK) return rc;
    h->pitems_measured = true;
  }
  h->last_kernel_ms = n_frames > 0 ? ms / float(n_frames) : 0.f;
  if (ms_out) *ms_out = ms;
  return MDWT_OK;
}

int batch_ready(mdwt_renderer* h) {
  int rc = ready_to_render(h);
  if (rc != MDWT_OK) return rc;
  if (h->views.empty()) return fail(h, MDWT_ERR_STATE, "no batched views (mdwt_batch_set_views)");
  HIPCHK(h, hipSetDevice(h->device));
  return MDWT_OK;
}

mdwt::CameraParams view_camera(const mdwt_view& v) {
  mdwt_renderer tmp;
  tmp.W = v.width;
  tmp.H = v.height;
  set_orientation(&tmp, F3{v.from[0], v.from[1], v.from[2]}, F3{v.at[0], v.at[1], v.at[2]},
                  F3{v.up[0], v.up[1], v.up[2]}, v.fovy);
  update_camera(&tmp);
  return tmp.cam;
}

// Block shape / tile changes: new single-view block order (the accumulated
// image is kept) and new batch items.
int reorder(mdwt_renderer* h) {
  if (h->inited) {
    HIPCHK(h, hipSetDevice(h->device));
    const uint32_t keep = h->accum_id;
    Float4* acc = h->d_accum;
    Float4* m2 = h->d_m2;
    h->d_accum = nullptr;
    h->d_m2 = nullptr;
    int rc = alloc_frame(h);
    if (rc != MDWT_OK) return rc;
    HIPCHK(h, hipMemcpy(h->d_accum, acc, h->fb_pixels * sizeof(Float4), hipMemcpyDeviceToDevice));
    if (m2 && h->d_m2)
      HIPCHK(h, hipMemcpy(h->d_m2, m2, h->fb_pixels * sizeof(Float4), hipMemcpyDeviceToDevice));
    (void)hipFree(acc);
    (void)hipFree(m2);
    h->accum_id = keep;
  }
  if (!h->views.empty()) return build_items(h);
  return MDWT_OK;
}

}  // namespace

extern "C" {

int mdwt_set_block(mdwt_renderer* h, int threads) {
  if (!h) return MDWT_ERR_ARG;
  int bw, bh;
  switch (threads) {
    case 64: bw = 8; bh = 8; break;
    case 128: bw = 16; bh = 8; break;
    case 256: bw = 16; bh = 16; break;
    default: return fail(h, MDWT_ERR_ARG, "workgroup size must be 64, 128 or 256");
  }
  if (bw == h->bw && bh == h->bh) return MDWT_OK;
  h->bw = bw;
  h->bh = bh;
  return reorder(h);
}

int mdwt_set_macrocells(mdwt_renderer* h, const int mc[3]) {
  if (!h || !mc) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  for (int i = 0; i < 3; ++i)
    if (mc[i] < 1) return fail(h, MDWT_ERR_ARG, "macrocell count must be positive");
  if (uint64_t(mc[0]) * mc[1] * mc[2] * uint64_t(h->C) > (1ull << 28))
    return fail(h, MDWT_ERR_LIMIT, "macrocell grid too large");
  std::vector<int> wx, wy, wz;
  if (!axis_windows(h->dims[0], mc[0], 0.f, h->remap.x, wx) ||
      !axis_windows(h->dims[1], mc[1], 0.f, h->remap.y, wy) ||
      !axis_windows(h->dims[2], mc[2], 0.f, h->remap.z, wz))
    return fail(h, MDWT_ERR_ARG, "macrocell projection windows are not intervals");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int i = 0; i < 3; ++i) h->mc[i] = mc[i];
  int rc = build_range_grid(h, wx, wy, wz);
  if (rc != MDWT_OK) return rc;
  reset_acc(h);
  if (!h->views.empty() && (rc = mdwt_batch_reset(h)) != MDWT_OK) return rc;
  return rebuild_majorants(h);
}

int mdwt_batch_set_views(mdwt_renderer* h, const mdwt_view* views, int n) {
  if (!h || !views) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if (n < 1 || n > 4096) return fail(h, MDWT_ERR_ARG, "1 to 4096 views");
  for (int i = 0; i < n; ++i) {
    const mdwt_view& v = views[i];
    if (v.width < 1 || v.height < 1 || v.width > 8192 || v.height > 8192)
      return fail(h, MDWT_ERR_ARG, "view size must be 1..8192");
    if (v.spp < 1 || v.spp > 1024) return fail(h, MDWT_ERR_ARG, "view spp must be 1..1024");
    if (!finite3(v.from) || !finite3(v.at) || !finite3(v.up) || !std::isfinite(v.fovy))
      return fail(h, MDWT_ERR_ARG, "non-finite view camera");
  }
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  free_batch(h);
  for (int i = 0; i < n; ++i) {
    const mdwt_view& v = views[i];
    mdwt::ViewParams vp{};
    vp.cam = view_camera(v);
    vp.width = v.width;
    vp.height = v.height;
    vp.spp = v.spp;
    vp.seed = v.seed;
    const size_t bytes = size_t(v.width) * v.height * sizeof(Float4);
    if (hipMalloc(&vp.accum, bytes) != hipSuccess ||
        (h->moments && hipMalloc(&vp.m2, bytes) != hipSuccess)) {
      h->views.push_back(vp);
      free_batch(h);
      return fail(h, MDWT_ERR_OOM, "view accumulators");
    }
    h->views.push_back(vp);
    HIPCHK(h, hipMemset(vp.accum, 0, bytes));
    if (vp.m2) HIPCHK(h, hipMemset(vp.m2, 0, bytes));
  }
  HIPCHK(h, hipMalloc(&h->d_views, h->views.size() * sizeof(mdwt::ViewParams)));
  HIPCHK(h, hipMemcpy(h->d_views, h->views.data(), h->views.size() * sizeof(mdwt::ViewParams),
                      hipMemcpyHostToDevice));
  HIPCHK(h, hipMalloc(&h->d_work, 4 * sizeof(unsigned)));
  const unsigned zero[4] = {0, 0, 0, 0};
  HIPCHK(h, hipMemcpy(h->d_work, zero, sizeof(zero), hipMemcpyHostToDevice));
  h->batch_frame = h->batch_accum = 0;
  return build_items(h);
}

int mdwt_batch_render(mdwt_renderer* h, int n_frames, float* ms_out) {
  if (!h) return MDWT_ERR_ARG;
  int rc = batch_ready(h);
  if (rc != MDWT_OK) return rc;
  if (n_frames < 0) return fail(h, MDWT_ERR_ARG, "negative frame count");
  switch (h->batch_launch) {
    case 1: return batch_render_one_launch(h, n_frames, ms_out);
    case 2: return batch_render_persistent(h, n_frames, ms_out);
    default: return batch_render_views(h, n_frames, ms_out);
  }
}

int mdwt_batch_count(mdwt_renderer* h, uint64_t counts[4]) {
  if (!h || !counts) return MDWT_ERR_ARG;
  int rc = batch_ready(h);
  if (rc != MDWT_OK) return rc;
  // the counting pass of form 2 is form 1's (the same per-item work)
  HIPCHK(h, hipMemsetAsync(h->d_counters, 0, 4 * sizeof(unsigned long long), h->stream));
  if (h->batch_launch >= 1)
    HIPCHK(h, mdwt::launch_views(views_params(h), h->mode, h->dtype, true, h->stream));
  else
    for (size_t i = 0; i < h->views.size(); ++i)
      HIPCHK(h, mdwt::launch_render(view_params(h, i), h->mode, h->dtype, true, h->stream));
  HIPCHK(h, hipMemcpyAsync(counts, h->d_counters, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return MDWT_OK;
}

int mdwt_batch_set_launch(mdwt_renderer* h, int launch) {
  if (!h) return MDWT_ERR_ARG;
  if (launch < 0 || launch > 2)
    return fail(h, MDWT_ERR_ARG,
                "batch launch must be 0 (a launch per view), 1 (one launch for all views) or 2 "
                "(one persistent launch, graph replay)");
  h->batch_launch = launch;
  return MDWT_OK;
}

int mdwt_batch_reset(mdwt_renderer* h) {
  if (!h) return MDWT_ERR_ARG;
  if (h->views.empty()) return fail(h, MDWT_ERR_STATE, "no batched views");
  h->batch_accum = 0;  // form 2 copies the ids to the device before its replays
  return MDWT_OK;
}

int mdwt_batch_read_view(mdwt_renderer* h, int view, float* rgba, float* m2) {
  if (!h) return MDWT_ERR_ARG;
  if (view < 0 || size_t(view) >= h->views.size()) return fail(h, MDWT_ERR_ARG, "no such view");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const auto& v = h->views[size_t(view)];
  const size_t bytes = size_t(v.width) * v.height * sizeof(Float4);
  if (m2 && !v.m2) return fail(h, MDWT_ERR_STATE, "second moments are off (mdwt_set_moments)");
  if (rgba) HIPCHK(h, hipMemcpy(rgba, v.accum, bytes, hipMemcpyDeviceToHost));
  if (m2) HIPCHK(h, hipMemcpy(m2, v.m2, bytes, hipMemcpyDeviceToHost));
  return MDWT_OK;
}

int mdwt_batch_view_times(mdwt_renderer* h, float* us) {
  if (!h || !us) return MDWT_ERR_ARG;
  if (h->views.empty()) return fail(h, MDWT_ERR_STATE, "no batched views");
  if (h->view_us.size() != h->views.size())
    return fail(h, MDWT_ERR_STATE, "no batched frame rendered yet");
  for (size_t k = 0; k < h->view_us.size(); ++k) us[k] = h->view_us[k];
  return MDWT_OK;
}

int mdwt_autotune(mdwt_renderer* h, const int* mc_sizes, int n_mc, const int* threads, int n_threads,
                  const int* tiles, int n_tiles, int frames, float* table, int best[3]) {
  if (!h || !mc_sizes || !threads || !tiles || n_mc < 1 || n_threads < 1 || n_tiles < 1)
    return MDWT_ERR_ARG;
  if (frames < 1) return fail(h, MDWT_ERR_ARG, "frames must be positive");
  const bool batch = !h->views.empty();
  int rc = batch ? batch_ready(h) : ready_to_render(h);
  if (rc != MDWT_OK) return rc;
  const int keep_order = h->order_mode;
  h->order_mode = 1;  // the tile order is what runs while timing
  float best_ms = std::numeric_limits<float>::max();
  int bi[3] = {mc_sizes[0], threads[0], tiles[0]};
  int row = 0;
  for (int a = 0; a < n_mc && rc == MDWT_OK; ++a) {
    const int s = mc_sizes[a];
    if (s < 1) {
      rc = fail(h, MDWT_ERR_ARG, "macrocell size must be positive");
      break;
    }
    const int mc[3] = {std::max(1, h->dims[0] / s), std::max(1, h->dims[1] / s),
                       std::max(1, h->dims[2] / s)};
    if ((rc = mdwt_set_macrocells(h, mc)) != MDWT_OK) break;
    uint64_t cnt[4] = {0, 0, 0, 0};
    for (int b = 0; b < n_threads && rc == MDWT_OK; ++b) {
      if ((rc = mdwt_set_block(h, threads[b])) != MDWT_OK) break;
      for (int c = 0; c < n_tiles && rc == MDWT_OK; ++c) {
        if ((rc = mdwt_set_tile_order(h, 1, tiles[c])) != MDWT_OK) break;
        if (batch) {
          if (b == 0 && c == 0 && (rc = mdwt_batch_count(h, cnt)) != MDWT_OK) break;
          if ((rc = mdwt_batch_render(h, 1, nullptr)) != MDWT_OK) break;
          if ((rc = mdwt_batch_render(h, frames, nullptr)) != MDWT_OK) break;
        } else {
          if (b == 0 && c == 0 && (rc = mdwt_count_frame(h, cnt)) != MDWT_OK) break;
          if ((rc = mdwt_render(h, 1, nullptr)) != MDWT_OK) break;
          if ((rc = mdwt_render(h, frames, nullptr)) != MDWT_OK) break;
        }
        const float ms = h->last_kernel_ms;
        if (table) {
          float* r = table + 6 * row;
          r[0] = float(s);
          r[1] = float(threads[b]);
          r[2] = float(tiles[c]);
          r[3] = ms;
          r[4] = ms > 0.f ? float(double(cnt[0]) / (double(ms) * 1e-3)) : 0.f;
          r[5] = ms > 0.f ? float(double(cnt[1]) / (double(ms) * 1e-3)) : 0.f;
        }
        ++row;
        if (ms < best_ms) {
          best_ms = ms;
          bi[0] = s;
          bi[1] = threads[b];
          bi[2] = tiles[c];
        }
      }
    }
  }
  h->order_mode = keep_order;
  if (rc != MDWT_OK) return rc;
  const int mc[3] = {std::max(1, h->dims[0] / bi[0]), std::max(1, h->dims[1] / bi[0]),
                     std::max(1, h->dims[2] / bi[0])};
  if ((rc = mdwt_set_macrocells(h, mc)) != MDWT_OK) return rc;
  if ((rc = mdwt_set_block(h, bi[1])) != MDWT_OK) return rc;
  if ((rc = mdwt_set_tile_order(h, keep_order, bi[2])) != MDWT_OK) return rc;
  if (batch && (rc = mdwt_batch_reset(h)) != MDWT_OK) return rc;
  reset_acc(h);
  if (best) {
    best[0] = bi[0];
    best[1] = bi[1];
    best[2] = bi[2];
  }
  return MDWT_OK;
}

}  // extern "C"

extern "C" {

int mdwt_set_aov(mdwt_renderer* h, int on) {
  if (!h) return MDWT_ERR_ARG;
  if (!h->inited) return fail(h, MDWT_ERR_STATE, "renderer not initialized");
  if ((on != 0) == h->aov_on) return MDWT_OK;
  h->aov_on = on != 0;
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int rc = alloc_frame(h);  // (re)allocates the buffers; accumulation restarts
  reset_acc(h);
  return rc;
}

int mdwt_set_aov_frames(mdwt_renderer* h, int frames) {
  if (!h || frames < 0) return MDWT_ERR_ARG;
  if (frames == h->aov_frames) return MDWT_OK;
  h->aov_frames = frames;
  reset_acc(h);
  return MDWT_OK;
}

int mdwt_read_aov(mdwt_renderer* h, float* out) {
  if (!h || !out) return MDWT_ERR_ARG;
  if (!h->aov_on || !h->d_aov) return fail(h, MDWT_ERR_STATE, "auxiliary buffers are off");
  HIPCHK(h, hipSetDevice(h->device));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const int ng = mdwt::aov_groups(h->C), nv = 2 + h->C;
  std::vector<float> tmp(size_t(4 * ng) * h->fb_pixels);
  HIPCHK(h, hipMemcpy(tmp.data(), h->d_aov, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t p = 0; p < h->fb_pixels; ++p)
    std::memcpy(out + size_t(nv) * p, tmp.data() + size_t(4 * ng) * p, size_t(nv) * sizeof(float));
  return MDWT_OK;
}

}  // extern "C"
