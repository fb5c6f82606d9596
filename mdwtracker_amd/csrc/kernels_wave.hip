// Wave-coherent Woodcock tracking for -m 0 (N-DDA) and -m 2 (syncDDA) on at
// most four channels (frame-kernel schedule 1).
//
// The one-thread-per-pixel kernel nests cell -> channel -> free-flight step
// loops per lane, so a wave64 runs every loop level as long as its slowest
// lane: most free-flight steps run with a third of the lanes, most sample
// evaluations with a few. Here a wave owns 64 rays (an 8x8 pixel quad) and
// compacts their work through LDS, in rounds:
//
//   1. walk:    each live ray walks its DDA up to K cells ahead and writes one
//               record (cell, entry/exit t, visit) per cell; the records of
//               all rays are packed into one list (K grows as rays finish, so
//               a round always holds about kRecords records);
//   2. enter:   the wave takes the list 64 records at a time: one lane per
//               record loads the cell's majorants, draws the first free-flight
//               step of every channel from one Philox block and decides
//               "leaves the cell" / "reaches the end of the root box" /
//               "needs a sample";
//   3. sample:  the (record, channel) steps that need a sample are pushed on
//               a per-wave LDS stack (ballot + mbcnt) and popped 64 at a time:
//               threshold, trilinear sample, transfer function, and on a null
//               collision the next step, pushed again while it stays inside;
//   4. resolve: each ray reads the outcomes of its records in cell order with
//               the reference's rules and either ends (hit, or the root box's
//               end) or continues in the next round.
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
//     first cell with a hit, the smallest t, ties to the higher channel.
// The hit's colour is recomputed from the same sample expression. Images are
// bit-identical to the one-thread-per-pixel kernel and the oracle; work
// counters come from the counting pass of that kernel (steps tracked past a
// ray's end here are speculation, not the reference's work).
#include "mdwt_internal.hpp"
#include "render_common.hpp"

namespace mdwt {
namespace {

constexpr int kRecords = 192;  // records per wave and round
constexpr int kMaxAhead = 16;  // cells a ray walks ahead per round, at most
constexpr int kStack = 128;    // sample stack: < 64 left + one channel's 64 pushes

enum : uint32_t { kLeave = 0u, kBoxEnd = 1u, kHit = 2u, kPending = 3u };

// Per-wave LDS. Record slot = k * n_live + rank for the k-th cell of the
// rank-th live ray, so the lanes of a walk or a resolve step touch
// consecutive words (no bank conflicts).
struct WaveLds {
  uint32_t cid[kRecords];
  float t0[kRecords], t1[kRecords];
  uint32_t visit[kRecords];
  uint16_t order[kRecords];  // packed list of the round's slots, ascending
  // first event of each (record, channel): 2-bit code per channel (bits 2c,
  // 2c + 1; pending entries are settled with LDS atomic AND, so lanes that
  // settle different channels of one record never overwrite each other) and
  // the event's t (grid units)
  uint32_t code[kRecords];
  float4 tev[kRecords];
  // sample stack: slot | channel << 8 | pair << 10, and the step's t
  uint32_t item[kStack];
  float item_t[kStack];
  // live rays of the round by rank, and every lane's ray
  uint8_t rank_lane[64];
  uint32_t pixel[64], mask[64];
  float w2g[64], g2w[64], tend[64];
  float ox[64], oy[64], oz[64], dx[64], dy[64], dz[64];
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
// settles channel c of a record from pending to code cd
__device__ __forceinline__ void settle(uint32_t* word, int c, uint32_t cd) {
  atomicAnd(word, ~((kPending ^ cd) << (2 * c)));
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

template <int DT>
__device__ __forceinline__ float4 cell_majorants(const Ctx<DT>& X, uint32_t cid) {
  if (X.C == 4) return *reinterpret_cast<const float4*>(X.maj + size_t(cid) * 4);
  const float* p = X.maj + size_t(cid) * X.C;
  return make_float4(p[0], X.C > 1 ? p[1] : 0.f, X.C > 2 ? p[2] : 0.f, 0.f);
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

// Pops up to 64 items of the sample stack (n_items) and tracks each one step
// further: null-collision test at the item's t, then the next free-flight step.
template <int DT>
__device__ __forceinline__ void sample_pass(const RenderParams& P, const Ctx<DT>& X, WaveLds& L,
                                            const float* s_a, uint32_t n_live, uint32_t& n_items,
                                            V3 w2u, uint32_t sample) {
  const int lane = threadIdx.x & 63;
  const uint32_t n = n_items < 64u ? n_items : 64u;
  const uint32_t base = n_items - n;
  const bool act = uint32_t(lane) < n;
  uint32_t it = 0u;
  float t = 0.f;
  if (act) {
    it = L.item[base + lane];
    t = L.item_t[base + lane];
  }
  wave_sync();
  bool push = false;
  float tn = 0.f;
  const uint32_t slot = it & 0xffu;
  const int c = int((it >> 8) & 3u);
  const uint32_t j = it >> 10;
  if (act) {
    const int rl = L.rank_lane[slot % n_live];
    const uint32_t cid = L.cid[slot];
    const float t1 = L.t1[slot];
    const uint32_t v = L.visit[slot];
    const float m = X.maj[size_t(cid) * X.C + c];
    const float w2g = L.w2g[rl], g2w = L.g2w[rl];
    uint32_t w[4];
    philox(P.fr.frame_id, sample, 1u + ((v << 4) | uint32_t(c)), (j >> 1) << 1,
           P.fr.seed, L.pixel[rl], w);
    const float thr = to_unit((j & 1u) ? w[2] : w[0]) * m;
    const float un = to_unit((j & 1u) ? w[3] : w[1]);
    const float tWorld = t * g2w;
    const V3 xT = (v3(L.ox[rl], L.oy[rl], L.oz[rl]) + tWorld * v3(L.dx[rl], L.dy[rl], L.dz[rl])) * w2u;
    const float value = X.sample(xT, c);
    bool hit = false;
    if (!(value != value)) {
      const float a = alpha_lookup(s_a, X.chtf[c], value);
      hit = a > 0.0f && thr < a;
    }
    if (hit) {
      reinterpret_cast<float*>(&L.tev[slot])[c] = t;
      settle(&L.code[slot], c, kHit);
    } else {
      tn = t - (plogf_poly(1.0f - un) / m) * X.dt * w2g;
      if (tn >= t1) {
        settle(&L.code[slot], c, kLeave);
      } else if (tn * g2w >= L.tend[rl]) {
        reinterpret_cast<float*>(&L.tev[slot])[c] = tn;
        settle(&L.code[slot], c, kBoxEnd);
      } else {
        push = true;
      }
    }
  }
  wave_sync();
  const uint64_t pm = __ballot(push);
  if (push) {
    const uint32_t pos = base + lane_rank(pm);
    L.item[pos] = slot | (uint32_t(c) << 8) | ((j + 1u) << 10);
    L.item_t[pos] = tn;
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

template <int MODE, int DT>
__device__ __forceinline__ void render_wave_body(const RenderParams& P) {
  extern __shared__ float s_alpha[];
  __shared__ ChanTf s_ch[kMaxChannels];
  __shared__ WaveLds s_wave[kThreads / 64];
  const uint64_t t_begin = P.fr.wave_cost ? wall_clock64() : 0ull;
  load_alpha_lds(P, s_alpha, s_ch);
  __syncthreads();  // the only workgroup barrier: the waves run independently from here

  Ctx<DT> X;
  // colours of a hit are looked up in the HBM copy of the tables
  X.init(P, reinterpret_cast<const float4*>(P.tf.tab), s_ch);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  WaveLds& L = s_wave[wave];
  const int C = X.C;

  const uint32_t slot_b = block_slot(blockIdx.x, uint32_t(P.fr.n_blocks));
  const uint32_t bxy = P.fr.block_order[slot_b];
  const int qs = P.fr.bw == 16 ? 1 : 0;
  const int px = int(bxy & 0xffffu) * P.fr.bw + (wave & qs) * 8 + (lane & 7);
  const int py = int(bxy >> 16) * P.fr.bh + (wave >> qs) * 8 + (lane >> 3);
  const bool valid = px < P.fr.width && py < P.fr.height;
  const uint32_t pixel = valid ? uint32_t(px) + uint32_t(P.fr.width) * uint32_t(py) : 0u;
  const V3 blo = load3(P.vol.blo), bhi = load3(P.vol.bhi);
  const V3 ext = bhi - blo;
  const V3 w2u = v3(1.f / ext.x, 1.f / ext.y, 1.f / ext.z);
  const uint32_t all_ch = (1u << C) - 1u;

  float fr = 0.f, fg = 0.f, fb = 0.f, fa = 1.f;
  for (int s = 0; s < P.fr.spp; ++s) {
    // ---- ray set-up (mainRG, deviceCode.cu:293-321; the program's grid ray)
    Dda d;
    bool alive = false;
    {
      Rng rng;
      rng.init(P.fr.seed, pixel, P.fr.frame_id, uint32_t(s), 0u);
      const float jit = rng.jitter();
      V3 org, dir;
      camera_ray(P.cam, px, py, P.fr.width, P.fr.height, jit, org, dir);
      float c0, c1;
      if (valid && clip_root(org, dir, blo, bhi, c0, c1)) {
        const GridRay g = make_grid_ray(org, dir, blo, bhi, X.mx, X.my, X.mz);
        alive = d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz);
        L.w2g[lane] = g.worldToGridT;
        L.g2w[lane] = g.gridToWorldT;
        L.tend[lane] = c1;
        L.ox[lane] = org.x;
        L.oy[lane] = org.y;
        L.oz[lane] = org.z;
        L.dx[lane] = dir.x;
        L.dy[lane] = dir.y;
        L.dz[lane] = dir.z;
      }
      L.pixel[lane] = pixel;
      L.mask[lane] = all_ch;
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
      const uint32_t rank = lane_rank(live);
      // ---- 1. walk: up to K cells per live ray
      uint32_t cnt = 0u;
      if (alive) {
        L.rank_lane[rank] = uint8_t(lane);
        L.mask[lane] = mask;
        if (!exhausted) {
          for (int k = 0; k < K; ++k) {
            float t0, t1;
            d.range(t0, t1);
            const uint32_t sl = uint32_t(k) * n_live + rank;
            L.cid[sl] = uint32_t(d.cx + d.cy * X.mx + d.cz * X.mx * X.my);
            L.t0[sl] = t0;
            L.t1[sl] = t1;
            L.visit[sl] = visit;
            ++visit;
            ++cnt;
            if (!d.advance()) {
              exhausted = true;
              break;
            }
          }
        }
      }
      // the round's slots, ascending (k-major)
      uint32_t total = 0u;
      for (int k = 0; k < K; ++k) {
        const uint64_t m = __ballot(uint32_t(k) < cnt);
        if (uint32_t(k) < cnt) L.order[total + lane_rank(m)] = uint16_t(uint32_t(k) * n_live + rank);
        total += uint32_t(__popcll(m));
      }
      wave_sync();
      // ---- 2. enter: first step of every channel of every record
      uint32_t n_items = 0u;
      for (uint32_t b = 0; b < total; b += 64u) {
        const uint32_t i = b + uint32_t(lane);
        const bool act = i < total;
        uint32_t sl = 0u;
        float4 tq = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t need = 0u;
        if (act) {
          sl = L.order[i];
          const int rl = L.rank_lane[sl % n_live];
          const uint32_t cid = L.cid[sl];
          const float t0 = L.t0[sl], t1 = L.t1[sl];
          const uint32_t v = L.visit[sl];
          const float w2g = L.w2g[rl], g2w = L.g2w[rl], tend = L.tend[rl];
          const uint32_t msk = L.mask[rl];
          const float4 m4 = cell_majorants(X, cid);
          uint32_t w[4];
          philox(P.fr.frame_id, uint32_t(s), 1u + ((v << 4) | 12u), 0u, P.fr.seed, L.pixel[rl], w);
          uint32_t codes = 0u;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            uint32_t cd = kLeave;
            const float m = comp(m4, c);
            if (c < C && m != 0.0f && ((msk >> c) & 1u)) {
              const float t = t0 - (plogf_poly(1.0f - to_unit(w[c])) / m) * X.dt * w2g;
              if (t < t1) {
                set_comp(tq, c, t);
                if (t * g2w >= tend) {
                  cd = kBoxEnd;
                } else {
                  cd = kPending;
                  need |= 1u << c;
                }
              }
            }
            codes |= cd << (2 * c);
          }
          L.code[sl] = codes;
          L.tev[sl] = tq;
        }
        wave_sync();
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bool nc = (need >> c) & 1u;
          const uint64_t pm = __ballot(nc);
          if (nc) {
            const uint32_t pos = n_items + lane_rank(pm);
            L.item[pos] = sl | (uint32_t(c) << 8);
            L.item_t[pos] = comp(tq, c);
          }
          n_items += uint32_t(__popcll(pm));
          wave_sync();
          // ---- 3. sample: full batches as soon as there are any
          while (n_items >= 64u)
            sample_pass<DT>(P, X, L, s_alpha, n_live, n_items, w2u, uint32_t(s));
        }
      }
      while (n_items > 0u) sample_pass<DT>(P, X, L, s_alpha, n_live, n_items, w2u, uint32_t(s));
      wave_sync();
      // ---- 4. resolve: the ray's records in cell order
      if (alive) {
        bool done = false;
        for (uint32_t k = 0; k < cnt && !done; ++k) {
          const uint32_t sl = k * n_live + rank;
          const uint32_t codes = L.code[sl];
          const float4 te = L.tev[sl];
          if (MODE == 2) {
            float minT = L.t1[sl];
            int hc = -1;
            bool term = false;
            for (int c = 0; c < C; ++c) {
              const uint32_t cd = (codes >> (2 * c)) & 3u;
              const float tc = comp(te, c);
              if (cd == kHit && tc < minT) {
                hc = c;
                minT = tc;
              } else if (cd == kBoxEnd && tc < minT && c == C - 1) {
                term = true;
              }
            }
            if (hc >= 0) {
              hit = true;
              hit_c = hc;
              hit_t = minT;
            }
            done = hc >= 0 || term;
          } else {
            float best = 1e20f;
            int hc = -1;
            for (int c = 0; c < C; ++c) {
              if (!((mask >> c) & 1u)) continue;
              const uint32_t cd = (codes >> (2 * c)) & 3u;
              const float tc = comp(te, c);
              if (cd == kHit && tc <= best) {
                hc = c;
                best = tc;
              } else if (cd == kBoxEnd) {
                mask &= ~(1u << c);
              }
            }
            if (hc >= 0) {
              hit = true;
              hit_c = hc;
              hit_t = best;
            }
            done = hc >= 0 || mask == 0u;
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
      const float g2w = L.g2w[lane];
      const float tWorld = hit_t * g2w;
      const V3 xT = (v3(L.ox[lane], L.oy[lane], L.oz[lane]) +
                     tWorld * v3(L.dx[lane], L.dy[lane], L.dz[lane])) * w2u;
      float r2;
      (void)X.tf_alpha(X.sample(xT, hit_c), hit_c, r2);
      X.tf_rgb(r2, hit_c, h);
      h.tHit = tWorld;
    }
    composite_sample(P, hit, h.r, h.g, h.b, h.a, 1.0f, 0.0f, 1.0f, h.channel, fr, fg, fb, fa);
    wave_sync();
  }
  if (valid) finish_pixel(P, P.fr.accum, P.fr.m2, P.fr.accum_id, P.fr.spp, pixel, fr, fg, fb, fa);
  if (P.fr.wave_cost && lane == 0)
    P.fr.wave_cost[(blockDim.x >> 6) * slot_b + uint32_t(wave)] = uint32_t(wall_clock64() - t_begin);
}

template <int MODE, int DT>
__global__ __launch_bounds__(256) void render_wave(RenderParams P) {
  render_wave_body<MODE, DT>(P);
}

template <int MODE>
RenderFn wave_fn(int dtype) {
  switch (dtype) {
    case MDWT_DT_U8: return render_wave<MODE, MDWT_DT_U8>;
    case MDWT_DT_U16: return render_wave<MODE, MDWT_DT_U16>;
    case MDWT_DT_F16: return render_wave<MODE, MDWT_DT_F16>;
    default: return render_wave<MODE, MDWT_DT_F32>;
  }
}

}  // namespace

bool wave_schedule_applies(const RenderParams& p, int mode, bool count, bool lit) {
  return (mode == 0 || mode == 2) && !count && !lit && p.vol.channels <= 4;
}

RenderFn select_wave_kernel(int mode, int dtype) {
  return mode == 0 ? wave_fn<0>(dtype) : wave_fn<2>(dtype);
}

size_t wave_lds_bytes(const RenderParams& p) {
  return size_t(p.vol.channels) * size_t(p.tf.stride) * sizeof(float);
}

}  // namespace mdwt
