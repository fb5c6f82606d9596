// Wave-coherent Woodcock tracking for the tracking modes, up to twelve
// channels: -m 0 (N-DDA), -m 1 (one clock per channel), -m 2 (syncDDA), -m 3
// (cumulative majorant) and -m 4 / 5 (blend tracking) (frame-kernel schedule 1).
//
// The one-thread-per-pixel kernel nests cell -> channel -> free-flight step
// loops per lane, so a wave64 runs every loop level as long as its slowest
// lane: most free-flight steps run with a third of the lanes, most sample
// evaluations with a few. Here a wave owns 64 rays (an 8x8 pixel quad) and
// compacts their work through LDS, in rounds:
//
//   1. walk:    each live ray walks its DDA up to K cells ahead and writes one
//               record (cell, entry/exit t, visit) per cell; the records of
//               all rays are packed into one dense, k-major list (K grows as
//               rays finish, so a round holds up to kRecords records);
//   2. enter:   the wave takes the list 64 records at a time: one lane per
//               record loads the cell's majorants and draws the first
//               free-flight distance of every channel from one Philox block;
//               a conservative bound settles the (record, channel) pairs
//               whose first step certainly leaves the cell, the others are
//               pushed on the item stack (ballot + mbcnt);
//   3. passes:  full batches of 64 items popped from the stack: a first step
//               is computed exactly and, when it lands inside the cell, gets
//               its null-collision test (threshold, trilinear sample,
//               transfer function); a rejected step's successor is pushed
//               back as a test item;
//   4. tail:    the round's last, partial batch runs each lane to its item's
//               end;
//   5. resolve: only events write: a hit, or the end of the root box, is an
//               LDS atomic minimum on the ray's 64-bit event key, so each ray
//               reads its first event of the round and either ends or
//               continues in the next round.
//
// Exactness. With the per-(visit, channel) random streams of mdwt_device.hpp,
// the steps of every (cell, channel) of a ray are a function of (visit,
// channel, step index) alone, so they can be computed in any order, by any
// lane, and cells can be tracked before it is known whether the ray gets
// there. Given the first event of each (cell, channel) - a hit at t, the end
// of the root box at t, or leaving the cell - the sequential programs'
// results follow exactly:
//   * altMultiMajDTCH (-m 2, deviceCode.cu:747-856): in the first cell where
//     some channel hits, or where the last channel reaches the box's end
//     before every hit of the lower channels, the ray ends; the hit is the
//     smallest t, ties to the lower channel (channel c tracks up to minT, the
//     nearest hit of channels < c, and stops at t >= minT);
//   * baseLineDTCH (-m 0, deviceCode.cu:858-970): every channel walks the
//     same cells with the same visit numbers; a channel's pass ends at its
//     first event (a box-end event removes the channel from later cells), and
//     a hit is kept when t <= the nearest hit so far, so the ray ends in the
//     first cell with a hit, the smallest t, ties to the higher channel;
//   * multiMajDTCH (-m 1, deviceCode.cu:619-745): the channels' clocks fire
//     in time order, the lower channel first on a tie, so the ray ends at the
//     earliest hit or clock past the box's end of any channel;
//   * cummilativeDTCH / blendDTCH (-m 3 / 4 / 5, deviceCode.cu:513-617,
//     972-1065): one summed majorant per cell, tracked as a single channel
//     whose null-collision test samples every channel; the first hit or box
//     end ends the ray.
// The hit's colour is recomputed from the same sample expression. Images are
// bit-identical to the one-thread-per-pixel kernel and the oracle; work
// counters come from the counting pass of that kernel (steps tracked past a
// ray's end here are speculation, not the reference's work).
#include "mdwt_internal.hpp"
#include "render_paths.hpp"  // blend_channels (-m 4 / 5)

namespace mdwt {
namespace {

constexpr int kRecords = 412;  // records per wave and round
constexpr int kMaxAhead = 16;  // cells a ray walks ahead per round, at most (-m 0: 64 / C above 4 channels)
constexpr int kStack = 128;    // the item stack: < 64 left + one channel's 64 pushes
// Safety factor of the certain-exit test (see the enter phase): far above the
// few fp32 roundings (relative 2^-21) that it has to absorb.
constexpr float kExitSafety = 1.0001f;
constexpr uint32_t kVisitBits = 26;  // visits of one ray below 2^26 (checked on the host)
constexpr uint32_t kVisitMask = (1u << kVisitBits) - 1u;
constexpr uint64_t kNoEvent = ~0ull;
// stack entry id: record slot (9 bits) | channel << 9 (4 bits) | first-step
// flag << 13 | pair << 14
constexpr uint32_t kSlotMask = 0x1ffu;
constexpr uint32_t kChanShift = 9;
constexpr uint32_t kFirstBit = 1u << 13;  // a first step (x = uniform), not a test (x = t)
constexpr uint32_t kPairShift = 14;

// A record: the macrocell of one (ray, visit), its entry and exit t (grid
// units), the visit number and the ray's lane (vl = visit | lane << 26), read
// and written as one 16-byte word.
struct Record {
  uint32_t cid;
  float t0, t1;
  uint32_t vl;
};
// A ray of the wave (one per lane): world origin and direction, grid <-> world
// t scales, end of the root box (world t), 1 / (dt * worldToGridT), pixel and
// the channels still tracked (-m 0).
struct WaveRay {
  float4 org_w2g;   // origin, worldToGridT
  float4 dir_g2w;   // direction, gridToWorldT
  float tend, idw;
  uint32_t pixel, mask;
};
// A stack entry: the id above, a t (a test) or a uniform (a first step), and
// the channel's majorant in the record's cell.
struct Item {
  uint32_t id;
  float x, m;
};
// The random streams of one sample of a view: frame id, sample index, seed.
struct Draw {
  uint32_t frame, sample, seed;
};

// Per-wave LDS. The records of a round are dense and k-major: the k-th cells
// of all rays that walk that far are consecutive, so the lanes of a walk step
// or an enter batch touch consecutive records. A ray's first event of the
// round (the smallest key among its hits and the box ends that end it, see
// event_key / clock_key) is an LDS atomic minimum, and -m 0's box ends set
// bits of a word, so only events write: leaving a cell writes nothing.
struct WaveLds {
  Record rec[kRecords];
  Item stk[kStack];  // first steps (x = uniform) and null-collision tests (x = t)
  WaveRay ray[64];
  unsigned long long ev[64];   // per ray: first event key of the round
  unsigned long long box[64];  // -m 0, per ray: end-of-box bits (see on_box_end)
};

__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}
// Orders the wave's LDS accesses across a phase boundary (lanes read what
// other lanes of the same wave wrote; the waves of a workgroup never share
// LDS, so no workgroup barrier is needed or allowed: they loop independently).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
// fp32 -> uint32 with the same order (negative values below positive ones)
__device__ __forceinline__ uint32_t order_bits(float t) {
  const uint32_t b = __float_as_uint(t);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_order_bits(uint32_t o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}
// Event key (-m 0 / 2 / 3 / 4 / 5): visit, then a flag (0: hit, 1: the end of
// the root box: -m 2's last channel, -m 3-5's summed majorant), then t, then
// 4 channel bits: 15 - channel for -m 0 (ties to the higher channel), the
// channel otherwise (ties to the lower one). Within a visit any hit ends
// -m 2's ray before the end of the box does.
__device__ __forceinline__ uint64_t event_key(uint32_t visit, uint32_t flag, float t, uint32_t ch) {
  return (uint64_t(visit) << 37) | (uint64_t(flag) << 36) | (uint64_t(order_bits(t)) << 4) | ch;
}
// -m 1: the channels' clocks fire in time order, the lower channel first on a
// tie, and any channel's clock past the box's end ends the ray: key visit,
// then t, then the channel, then the flag.
__device__ __forceinline__ uint64_t clock_key(uint32_t visit, uint32_t flag, float t, uint32_t ch) {
  return (uint64_t(visit) << 37) | (uint64_t(order_bits(t)) << 5) | (uint64_t(ch) << 1) | flag;
}
template <int MODE>
__device__ __forceinline__ void on_hit(WaveLds& L, const Record& R, int c, float t) {
  if (MODE == 1) {
    atomicMin(&L.ev[R.vl >> kVisitBits], clock_key(R.vl & kVisitMask, 0u, t, uint32_t(c)));
    return;
  }
  const uint32_t ch = MODE == 0 ? uint32_t(15 - c) : uint32_t(c);
  atomicMin(&L.ev[R.vl >> kVisitBits], event_key(R.vl & kVisitMask, 0u, t, ch));
}
template <int MODE>
__device__ __forceinline__ void on_box_end(WaveLds& L, const Record& R, int c, int C, float t) {
  if (MODE == 1) {
    atomicMin(&L.ev[R.vl >> kVisitBits], clock_key(R.vl & kVisitMask, 1u, t, uint32_t(c)));
  } else if (MODE >= 3) {
    atomicMin(&L.ev[R.vl >> kVisitBits], event_key(R.vl & kVisitMask, 1u, t, 0u));
  } else if (MODE == 2) {
    if (c == C - 1) atomicMin(&L.ev[R.vl >> kVisitBits], event_key(R.vl & kVisitMask, 1u, t, 0u));
  } else {
    // bit (visit mod 16) * 4 + c; above four channels, (visit mod kv) * C + c
    // with kv = 64 / C visits in the word
    const uint32_t v = R.vl & kVisitMask;
    const uint32_t bit = C <= 4 ? (((v & 15u) << 2) | uint32_t(c))
                                : (v % uint32_t(64 / C)) * uint32_t(C) + uint32_t(c);
    atomicOr(&L.box[R.vl >> kVisitBits], 1ull << bit);
  }
}
__device__ __forceinline__ float comp(const float4& v, int c) {
  return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}
__device__ __forceinline__ void set_comp(float4& v, int c, float x) {
  if (c == 0) v.x = x;
  else if (c == 1) v.y = x;
  else if (c == 2) v.z = x;
  else v.w = x;
}

// The majorants of channels 4g .. 4g + 3 in cell cid (0 past the last channel).
template <int DT>
__device__ __forceinline__ float4 cell_majorants(const Ctx<DT>& X, uint32_t cid, int g) {
  if (X.C == 4) return *reinterpret_cast<const float4*>(X.maj + size_t(cid) * 4);
  const int c0 = 4 * g;
  const float* p = X.maj + size_t(cid) * X.C + c0;
  return make_float4(p[0], X.C > c0 + 1 ? p[1] : 0.f, X.C > c0 + 2 ? p[2] : 0.f,
                     X.C > c0 + 3 ? p[3] : 0.f);
}

// transferFunction's opacity (deviceCode.cu:27-42) from the LDS table of
// opacities: the same operations as Ctx::tf_alpha on the float4 table.
__device__ __forceinline__ float alpha_lookup(const float* s_a, const ChanTf& t, float f) {
  if (f < t.dlo || f > t.dhi) return 0.0f;
  const float r1 = (f - t.dlo) * t.idw;
  const float r2 = (r1 - t.xlo) * t.ixw;
  const float x = r2 * float(t.n) - 0.5f;
  const float fl = floorf(x);
  const float fr = x - fl;
  const float a = s_a[t.off + clamp_index(fl, t.n)];
  const float b = s_a[t.off + clamp_index_plus1(fl, t.n)];
  return (a + fr * (b - a)) * t.os;
}

// The null-collision test of a step that landed at t inside its cell (before
// the root box's end), then the next free-flight step; pair j of the
// (visit, channel) stream. Returns true when the step needs another test (tn
// inside the cell and before the box's end); otherwise its event, if any, is
// recorded here.
template <int MODE, int DT>
__device__ __forceinline__ bool null_test(const RenderParams& P, const Ctx<DT>& X, WaveLds& L,
                                          const float* s_a, V3 w2u, const Draw& D, int c,
                                          const Record& R, const WaveRay& Y, uint32_t j, float t,
                                          float m, float& tn) {
  const float4 ow = Y.org_w2g, dg = Y.dir_g2w;
  uint32_t w[4];
  philox(D.frame, D.sample, 1u + (((R.vl & kVisitMask) << 4) | uint32_t(c)), (j >> 1) << 1, D.seed,
         Y.pixel, w);
  const float thr = to_unit((j & 1u) ? w[2] : w[0]) * m;
  const float un = to_unit((j & 1u) ? w[3] : w[1]);
  const float tWorld = t * dg.w;
  const V3 xT = (v3(ow.x, ow.y, ow.z) + tWorld * v3(dg.x, dg.y, dg.z)) * w2u;
  if (MODE == 4 || MODE == 5) {
    // blendDTCH (deviceCode.cu:972-1065): one blended sample of all channels
    const float4 sb = blend_channels<MODE, DT>(X, xT);
    if (sb.w > 0.0f && thr < sb.w) {
      on_hit<MODE>(L, R, 0, t);
      return false;
    }
  } else if (MODE == 3) {
    // cummilativeDTCH (deviceCode.cu:585-609): the channels in order take
    // their opacity off the threshold until one exceeds what is left
    float th = thr;
    for (int k = 0; k < X.C; ++k) {
      const float value = X.sample(xT, k);
      if (value != value) continue;
      const float a = alpha_lookup(s_a, X.chtf[k], value);
      if (a > 0.0f && th < a) {
        on_hit<MODE>(L, R, k, t);
        return false;
      }
      th -= a;
    }
  } else {
    const float value = X.sample(xT, c);
    bool hit = false;
    if (!(value != value)) {
      const float a = alpha_lookup(s_a, X.chtf[c], value);
      hit = a > 0.0f && thr < a;
    }
    if (hit) {
      on_hit<MODE>(L, R, c, t);
      return false;
    }
  }
  tn = t - (plogf_poly(1.0f - un) / m) * X.dt * ow.w;
  if (tn >= R.t1) return false;  // leaves the cell
  if (tn * dg.w >= Y.tend) {
    on_box_end<MODE>(L, R, c, X.C, tn);
    return false;
  }
  return true;
}

// Pops up to 64 items of the stack (n_items): a first step is computed
// exactly, as the sequential loop does, and when it lands inside the cell
// before the box's end gets its null-collision test; a test item is tested at
// its t. Full passes push a rejected step's successor back; the round's last
// pass (DRAIN, fewer than 64 items) keeps every lane on its item until it is
// settled.
template <int MODE, int DT, bool DRAIN>
__device__ __forceinline__ void item_pass(const RenderParams& P, const Ctx<DT>& X, WaveLds& L,
                                          const float* s_a, uint32_t& n_items, V3 w2u,
                                          const Draw& D) {
  const int lane = threadIdx.x & 63;
  const uint32_t n = n_items < 64u ? n_items : 64u;
  const uint32_t base = n_items - n;
  const bool act = uint32_t(lane) < n;
  Item it = {0u, 0.f, 0.f};
  if (act) it = L.stk[base + lane];
  wave_sync();
  bool push = false;
  float tn = 0.f;
  uint32_t j = it.id >> kPairShift;
  if (act) {
    const Record R = L.rec[it.id & kSlotMask];
    const int c = int((it.id >> kChanShift) & 15u);
    const WaveRay Y = L.ray[R.vl >> kVisitBits];
    float t = it.x;
    bool go = true;
    if (it.id & kFirstBit) {
      t = R.t0 - (plogf_poly(1.0f - it.x) / it.m) * X.dt * Y.org_w2g.w;
      go = false;
      if (t < R.t1) {  // otherwise it leaves the cell
        if (t * Y.dir_g2w.w >= Y.tend)
          on_box_end<MODE>(L, R, c, X.C, t);
        else
          go = true;
      }
    }
    while (go && null_test<MODE, DT>(P, X, L, s_a, w2u, D, c, R, Y, j, t, it.m, tn)) {
      ++j;
      if (!DRAIN) {
        push = true;
        break;
      }
      t = tn;
    }
  }
  wave_sync();
  if (DRAIN) {
    n_items = base;
    return;
  }
  const uint64_t pm = __ballot(push);
  if (push) {
    Item o;
    o.id = (it.id & (kFirstBit - 1u)) | (j << kPairShift);
    o.x = tn;
    o.m = it.m;
    L.stk[base + lane_rank(pm)] = o;
  }
  wave_sync();
  n_items = base + uint32_t(__popcll(pm));
}

// Transfer-function opacities and per-channel parameters -> LDS (the caller
// synchronises the workgroup afterwards). Colours stay in HBM: only a ray's
// hit reads them.
__device__ __forceinline__ void load_alpha_lds(const RenderParams& P, float* s_a, ChanTf* s_ch) {
  const int C = P.vol.channels;
  const float4* g = reinterpret_cast<const float4*>(P.tf.tab);
  const int total = C * P.tf.stride;
  for (int i = threadIdx.x; i < total; i += blockDim.x) s_a[i] = g[i].w;
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

// The 8x8 pixel quad of this wave in pixel block (bx, by) of view V (a
// workgroup's bw x bh block has one quad per wave). WIDE: more than four
// channels (-m 2), drawn in groups of four; a separate instantiation, since
// the group loop's code in the four-channel kernel cost it 3%.
template <int MODE, int DT, bool WIDE>
__device__ __forceinline__ void wave_block(const RenderParams& P, const Ctx<DT>& X, WaveLds& L,
                                           const float* s_alpha, const ViewCtx& V, int bx, int by) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int C = X.C;
  const int qs = P.fr.bw == 16 ? 1 : 0;
  const int px = bx * P.fr.bw + (wave & qs) * 8 + (lane & 7);
  const int py = by * P.fr.bh + (wave >> qs) * 8 + (lane >> 3);
  const bool valid = px < V.W && py < V.H;
  const uint32_t pixel = valid ? uint32_t(px) + uint32_t(V.W) * uint32_t(py) : 0u;
  const V3 blo = load3(P.vol.blo), bhi = load3(P.vol.bhi);
  const V3 ext = bhi - blo;
  const V3 w2u = v3(1.f / ext.x, 1.f / ext.y, 1.f / ext.z);
  const uint32_t all_ch = (1u << C) - 1u;

  float fr = 0.f, fg = 0.f, fb = 0.f, fa = 1.f;
  for (int s = 0; s < V.spp; ++s) {
    const Draw D = {V.frame_id, uint32_t(s), V.seed};
    // ---- ray set-up (mainRG, deviceCode.cu:293-321; the program's grid ray)
    Dda d;
    bool alive = false;
    float idw = 0.f;
    {
      Rng rng;
      rng.init(V.seed, pixel, V.frame_id, uint32_t(s), 0u);
      float jx, jy;
      rng.jitter2(P.fr.jitter2d != 0, jx, jy);
      V3 org, dir;
      camera_ray(*V.cam, px, py, V.W, V.H, jx, jy, org, dir);
      float c0, c1;
      WaveRay Y;
      Y.org_w2g = make_float4(org.x, org.y, org.z, 0.f);
      Y.dir_g2w = make_float4(dir.x, dir.y, dir.z, 0.f);
      Y.tend = 0.f;
      if (valid && clip_root(org, dir, blo, bhi, c0, c1)) {
        const GridRay g = make_grid_ray(org, dir, blo, bhi, X.mx, X.my, X.mz);
        alive = d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz);
        Y.org_w2g.w = g.worldToGridT;
        Y.dir_g2w.w = g.gridToWorldT;
        Y.tend = c1;
        idw = 1.0f / (X.dt * g.worldToGridT);
      }
      Y.idw = idw;
      Y.pixel = pixel;
      Y.mask = all_ch;
      L.ray[lane] = Y;
    }
    bool exhausted = false;
    uint32_t visit = 0u, mask = all_ch;
    bool hit = false;
    int hit_c = 0;
    float hit_t = 0.f;
    while (true) {
      const uint64_t live = __ballot(alive);
      if (live == 0ull) break;
      const uint32_t n_live = uint32_t(__popcll(live));
      int K = int(uint32_t(kRecords) / n_live);
      K = K > kMaxAhead ? kMaxAhead : K;
      if (MODE == 0 && WIDE && K > 64 / C) K = 64 / C;  // the end-of-box word's visits
      const uint32_t visit0 = visit;
      if (alive) {
        L.ray[lane].mask = mask;
        L.ev[lane] = kNoEvent;
        L.box[lane] = 0ull;
      }
      // ---- 1. walk: up to K cells per live ray, records dense in k-major order
      bool walking = alive && !exhausted;
      uint32_t total = 0u;
      for (int k = 0; k < K; ++k) {
        const uint64_t wm = __ballot(walking);
        if (wm == 0ull) break;
        if (walking) {
          float t0, t1;
          d.range(t0, t1);
          Record R;
          R.cid = uint32_t(d.cx + d.cy * X.mx + d.cz * X.mx * X.my);
          R.t0 = t0;
          R.t1 = t1;
          R.vl = visit | (uint32_t(lane) << kVisitBits);
          L.rec[total + lane_rank(wm)] = R;
          ++visit;
          if (!d.advance()) {
            exhausted = true;
            walking = false;
          }
        }
        total += uint32_t(__popcll(wm));
      }
      wave_sync();
      // ---- 2. enter: first step of every channel of every record. The
      // sequential loop's first step leaves the cell when
      //   t0 - (log(1 - u) / m) dt w2g >= t1   (fp32, each operation rounded).
      // Since -log(1 - u) >= u, the portable log is within 2^-22 of log, and
      // each rounding moves a value by at most 2^-24 relative, the step
      // certainly leaves when
      //   u >= ((t1 - t0) + t1 2^-23) m / (dt w2g) (1 + 2^-21),
      // which the test below over-estimates by the safety factor. Most steps
      // pass it with a few instructions; the rest are computed exactly, in
      // full batches, from the first-step stack.
      uint32_t n_items = 0u;
      const int groups = WIDE ? (C + 3) >> 2 : 1;  // channels in groups of 4: one Philox block each
      // -m 3 / 4 / 5 (one summed majorant per cell): each batch's record and
      // majorant are read one batch ahead, so their latency overlaps the batch
      // before (records do not change during the round). 3 % faster there; the
      // same for the per-channel majorants of -m 0 / 1 / 2 measured 1 % slower.
      constexpr bool kAhead = !WIDE && MODE >= 3;
      Record r_nx = {0u, 0.f, 0.f, 0u};
      float m_nx = 0.f;
      auto fetch_next = [&](const uint32_t i) {
        if (kAhead && i < total) {
          r_nx = L.rec[i];
          m_nx = X.maj[r_nx.cid];
        }
      };
      fetch_next(uint32_t(lane));
      for (uint32_t b = 0; b < total; b += 64u) {
        const uint32_t i = b + uint32_t(lane);
        const bool act = i < total;
        const Record r_cur = r_nx;
        const float m_cur = m_nx;
        fetch_next(i + 64u);
        // the first steps of channels 4g .. 4g + 3 of record i
        auto enter_group = [&](const int g) {
          float4 uq = make_float4(0.f, 0.f, 0.f, 0.f), m4 = uq;
          uint32_t need = 0u;
          if (act) {
            const Record R = kAhead ? r_cur : L.rec[i];
            const WaveRay& Y = L.ray[R.vl >> kVisitBits];
            const float q = ((R.t1 - R.t0) + R.t1 * 0x1p-23f) * Y.idw * kExitSafety;
            if (MODE >= 3) {
              // one summed majorant per cell, tracked at channel 0's positions
              const float m = kAhead ? m_cur : X.maj[R.cid];
              if (m != 0.0f) {
                uint32_t w[4];
                philox(D.frame, D.sample, 1u + (((R.vl & kVisitMask) << 4) | 12u), 0u, D.seed,
                       Y.pixel, w);
                const float u = to_unit(w[0]);
                m4.x = m;
                if (!(u >= q * m)) {
                  uq.x = u;
                  need = 1u;
                }
              }
            } else {
              m4 = cell_majorants(X, R.cid, g);
              uint32_t w[4];
              philox(D.frame, D.sample, 1u + (((R.vl & kVisitMask) << 4) | (12u + uint32_t(g))),
                     0u, D.seed, Y.pixel, w);
              const uint32_t msk = Y.mask >> (4 * g);
#pragma unroll
              for (int k = 0; k < 4; ++k) {
                const float m = comp(m4, k);
                if (4 * g + k < C && m != 0.0f && ((msk >> k) & 1u)) {
                  const float u = to_unit(w[k]);
                  if (!(u >= q * m)) {
                    set_comp(uq, k, u);
                    need |= 1u << k;
                  }
                }
              }
            }
          }
          constexpr int kPush = MODE >= 3 ? 1 : 4;  // tracked channels per record
#pragma unroll
          for (int k = 0; k < kPush; ++k) {
            const bool nc = (need >> k) & 1u;
            const uint64_t pm = __ballot(nc);
            if (nc) {
              Item e;
              e.id = i | (uint32_t(4 * g + k) << kChanShift) | kFirstBit;
              e.x = comp(uq, k);
              e.m = comp(m4, k);
              L.stk[n_items + lane_rank(pm)] = e;
            }
            n_items += uint32_t(__popcll(pm));
            // full batches as soon as there are any
            if (n_items >= 64u) {
              wave_sync();
              do {
                item_pass<MODE, DT, false>(P, X, L, s_alpha, n_items, w2u, D);
              } while (n_items >= 64u);
            }
          }
        };
        if (WIDE) {
          for (int g = 0; g < groups; ++g) enter_group(g);
        } else {
          enter_group(0);
        }
      }
      wave_sync();
      // fewer than 64 items left: one looping pass
      if (n_items > 0u) item_pass<MODE, DT, true>(P, X, L, s_alpha, n_items, w2u, D);
      wave_sync();
      // ---- 4. resolve: the ray's first event of the round
      if (alive) {
        const uint64_t e = L.ev[lane];
        bool done = false;
        if (MODE == 1) {
          // the earliest event of any channel's clock: a hit, or its clock
          // past the box's end
          if (e != kNoEvent) {
            done = true;
            if (!(e & 1ull)) {
              hit = true;
              hit_c = int((e >> 1) & 15ull);
              hit_t = from_order_bits(uint32_t(e >> 5));
            }
          }
        } else if (MODE >= 2) {
          // a hit in the first visit with an event (smallest t, ties to the
          // lower channel; -m 3 has at most one per visit), or else the end
          // of the box there (-m 2: the last channel's)
          if (e != kNoEvent) {
            done = true;
            if (!((e >> 36) & 1ull)) {
              hit = true;
              hit_c = int(e & 15ull);
              hit_t = from_order_bits(uint32_t(e >> 4));
            }
          }
        } else {
          // channel c leaves the ray at its first end-of-box visit; the ray
          // ends at its first hit (smallest t, ties to the higher channel)
          // unless every channel it still tracks left before that visit
          const uint64_t bx0 = L.box[lane];
          const uint32_t k_hit = e != kNoEvent ? uint32_t(e >> 37) - visit0 : 0xffffffffu;
          uint32_t k_all = 0u, left = 0u;
          bool all = true;
          if (!WIDE) {
            const uint32_t rot = (visit0 & 15u) << 2;
            const uint64_t bx = rot ? ((bx0 >> rot) | (bx0 << (64u - rot))) : bx0;  // bit 4k + c
            for (int c = 0; c < C; ++c) {
              if (!((mask >> c) & 1u)) continue;
              const uint64_t cb = bx & (0x1111111111111111ull << c);
              if (cb == 0ull) {
                all = false;
              } else {
                const uint32_t kc = uint32_t(__builtin_ctzll(cb)) >> 2;
                k_all = kc > k_all ? kc : k_all;
                left |= 1u << c;
              }
            }
          } else {
            // a ring of kv = 64 / C visits of C bits; the round's k-th visit
            // sits at slot (visit0 + k) mod kv
            const uint32_t kv = uint32_t(64 / C), cmask = (1u << C) - 1u;
            const uint32_t s0 = visit0 % kv;
            for (uint32_t k = 0; k < uint32_t(K); ++k) {
              const uint32_t slot = (s0 + k) % kv;
              const uint32_t chunk = uint32_t(bx0 >> (slot * uint32_t(C))) & cmask & mask;
              const uint32_t fresh = chunk & ~left;
              if (fresh) {
                left |= fresh;
                k_all = k;
              }
            }
            all = left == mask;
          }
          if (e != kNoEvent && (!all || k_hit <= k_all)) {
            done = true;
            hit = true;
            hit_c = 15 - int(e & 15ull);
            hit_t = from_order_bits(uint32_t(e >> 4));
          } else if (all) {
            done = true;
          } else {
            mask &= ~left;
          }
        }
        if (done || exhausted) alive = false;
      }
      wave_sync();
    }
    // ---- shade and composite the sample (deviceCode.cu:411-433)
    HitRec h;
    h.missed = !hit;
    h.a = 1.f;
    h.channel = hit_c;
    h.r = h.g = h.b = 0.f;
    if (hit) {
      const WaveRay& Y = L.ray[lane];
      const float tWorld = hit_t * Y.dir_g2w.w;
      const V3 xT = (v3(Y.org_w2g.x, Y.org_w2g.y, Y.org_w2g.z) +
                     tWorld * v3(Y.dir_g2w.x, Y.dir_g2w.y, Y.dir_g2w.z)) * w2u;
      if (MODE == 4 || MODE == 5) {
        // the blended sample's colour; the hit keeps channel 0
        const float4 sb = blend_channels<MODE, DT>(X, xT);
        h.r = sb.x;
        h.g = sb.y;
        h.b = sb.z;
      } else {
        float r2;
        (void)X.tf_alpha(X.sample(xT, hit_c), hit_c, r2);
        X.tf_rgb(r2, hit_c, h);
      }
      h.tHit = tWorld;
    }
    composite_sample(P, hit, h.r, h.g, h.b, h.a, 1.0f, 0.0f, 1.0f, h.channel, fr, fg, fb, fa);
    wave_sync();
  }
  if (valid) finish_pixel(P, V.accum, V.m2, V.accum_id, V.spp, pixel, fr, fg, fb, fa);
}

// Per-workgroup set-up: the opacity table and channel parameters in LDS (the
// only workgroup barrier: the waves run independently from here) and the
// per-wave LDS. The per-wave LDS is static, sized for a 256-thread block (with
// a dynamic array behind the opacity table the frame ran 15% slower); smaller
// blocks leave part of it unused.
#define MDWT_WAVE_PROLOGUE                                                        \
  extern __shared__ float s_alpha[];                                              \
  __shared__ ChanTf s_ch[kMaxChannels];                                           \
  __shared__ WaveLds s_wave[kThreads / 64];                                       \
  load_alpha_lds(P, s_alpha, s_ch);                                               \
  __syncthreads();                                                                \
  Ctx<DT> X;                                                                      \
  /* colours of a hit are looked up in the HBM copy of the tables */            \
  X.init(P, reinterpret_cast<const float4*>(P.tf.tab), s_ch);                     \
  WaveLds& L = s_wave[threadIdx.x >> 6];

// One view: workgroup i renders the pixel block at position block_slot(i) of
// the frame's block order.
template <int MODE, int DT, bool WIDE>
__global__ __launch_bounds__(256) void render_wave(RenderParams P) {
  const uint64_t t_begin = P.fr.wave_cost ? wall_clock64() : 0ull;
  MDWT_WAVE_PROLOGUE
  const uint32_t slot_b = block_slot(blockIdx.x, uint32_t(P.fr.n_blocks));
  const uint32_t bxy = P.fr.block_order[slot_b];
  ViewCtx V;
  V.cam = &P.cam;
  V.W = P.fr.width;
  V.H = P.fr.height;
  V.spp = P.fr.spp;
  V.seed = P.fr.seed;
  V.frame_id = P.fr.frame_id;
  V.accum_id = P.fr.accum_id;
  V.accum = P.fr.accum;
  V.m2 = P.fr.m2;
  wave_block<MODE, DT, WIDE>(P, X, L, s_alpha, V, int(bxy & 0xffffu), int(bxy >> 16));
  if (P.fr.wave_cost && (threadIdx.x & 63) == 0)
    P.fr.wave_cost[(blockDim.x >> 6) * slot_b + (threadIdx.x >> 6)] =
        uint32_t(wall_clock64() - t_begin);
}

__device__ __forceinline__ ViewCtx batch_view(const RenderParams& P, uint32_t code,
                                              uint32_t frame_id, uint32_t accum_id) {
  const ViewParams* vp = P.fr.views + (code >> 20);
  ViewCtx V;
  V.cam = &vp->cam;
  V.W = vp->width;
  V.H = vp->height;
  V.spp = vp->spp;
  V.seed = vp->seed;
  V.frame_id = frame_id;
  V.accum_id = accum_id;
  V.accum = vp->accum;
  V.m2 = vp->m2;
  return V;
}

// Batched views, launch form 1: workgroup i renders item i (view << 20 |
// by << 10 | bx) at the launch's frame and accumulation ids; with item_ticks
// set it records its start and end wall clock.
template <int MODE, int DT, bool WIDE>
__global__ __launch_bounds__(256) void render_views_wave(RenderParams P) {
  const uint64_t t_begin = P.fr.item_ticks ? wall_clock64() : 0ull;
  MDWT_WAVE_PROLOGUE
  const uint32_t code = P.fr.items[blockIdx.x];
  const ViewCtx V = batch_view(P, code, P.fr.frame_id, P.fr.accum_id);
  wave_block<MODE, DT, WIDE>(P, X, L, s_alpha, V, int(code & 0x3ffu), int((code >> 10) & 0x3ffu));
  if (P.fr.item_ticks) {
    __syncthreads();
    if (threadIdx.x == 0) {
      P.fr.item_ticks[2 * blockIdx.x] = t_begin;
      P.fr.item_ticks[2 * blockIdx.x + 1] = wall_clock64();
    }
  }
}

// Batched views, launch form 2 (persistent): workgroup g renders items g,
// g + G, ... at the frame and accumulation ids in work[1] / work[2]; each wave
// moves to its next item as soon as its quad is done. Wave 0 records each
// item's start and end wall clock.
template <int MODE, int DT, bool WIDE>
__global__ __launch_bounds__(256) void render_views_wave_persistent(RenderParams P) {
  MDWT_WAVE_PROLOGUE
  const uint32_t fid = P.fr.work[1], aid = P.fr.work[2];
  const uint32_t n = uint32_t(P.fr.n_items);
  for (uint32_t it = blockIdx.x; it < n; it += gridDim.x) {
    const uint64_t t_item = P.fr.item_ticks ? wall_clock64() : 0ull;
    const uint32_t code = P.fr.items[it];
    const ViewCtx V = batch_view(P, code, fid, aid);
    wave_block<MODE, DT, WIDE>(P, X, L, s_alpha, V, int(code & 0x3ffu), int((code >> 10) & 0x3ffu));
    if (P.fr.item_ticks && threadIdx.x == 0) {
      P.fr.item_ticks[2 * it] = t_item;
      P.fr.item_ticks[2 * it + 1] = wall_clock64();
    }
  }
}
#undef MDWT_WAVE_PROLOGUE

// form: 0 one view, 1 batched views (one launch), 2 batched views (persistent)
template <int MODE, int DT, bool WIDE>
RenderFn wave_fn_form(int form) {
  return form == 0 ? render_wave<MODE, DT, WIDE>
                   : (form == 1 ? render_views_wave<MODE, DT, WIDE>
                                : render_views_wave_persistent<MODE, DT, WIDE>);
}
template <int MODE, bool WIDE>
RenderFn wave_fn(int dtype, int form) {
  switch (dtype) {
    case MDWT_DT_U8: return wave_fn_form<MODE, MDWT_DT_U8, WIDE>(form);
    case MDWT_DT_U16: return wave_fn_form<MODE, MDWT_DT_U16, WIDE>(form);
    case MDWT_DT_F16: return wave_fn_form<MODE, MDWT_DT_F16, WIDE>(form);
    default: return wave_fn_form<MODE, MDWT_DT_F32, WIDE>(form);
  }
}

}  // namespace

bool wave_schedule_applies(const RenderParams& p, int mode, bool count, bool lit) {
  // a ray visits at most mx + my + mz cells: visit numbers fit the records' 26 bits
  const uint64_t max_visits = uint64_t(p.mc[0]) + uint64_t(p.mc[1]) + uint64_t(p.mc[2]);
  return mode >= 0 && mode <= 5 && !count && !lit && p.vol.channels <= kMaxChannels &&
         max_visits < (uint64_t(1) << kVisitBits);
}

RenderFn select_wave_kernel(const RenderParams& p, int mode, int dtype, int form) {
  if (mode == 0)
    return p.vol.channels > 4 ? wave_fn<0, true>(dtype, form) : wave_fn<0, false>(dtype, form);
  if (mode == 3) return wave_fn<3, false>(dtype, form);
  if (mode == 4) return wave_fn<4, false>(dtype, form);
  if (mode == 5) return wave_fn<5, false>(dtype, form);
  if (mode == 1)
    return p.vol.channels > 4 ? wave_fn<1, true>(dtype, form) : wave_fn<1, false>(dtype, form);
  return p.vol.channels > 4 ? wave_fn<2, true>(dtype, form) : wave_fn<2, false>(dtype, form);
}

size_t wave_lds_bytes(const RenderParams& p) {
  static_assert(sizeof(WaveLds) % 16 == 0, "the opacity table follows the per-wave LDS");
  return size_t(p.vol.channels) * size_t(p.tf.stride) * sizeof(float);
}

}  // namespace mdwt
