// Kernels (b)+(c), flat-loop schedules.
//
// The one-thread-per-pixel kernel walks cell -> channel -> free-flight step
// in nested loops; in a wave64 the nesting makes every loop level run as long
// as its slowest lane, so most lanes idle through most steps. Here each lane
// is a small state machine (step / enter cell / next pass / start sample /
// finish sample) and the wave runs one flat loop: every iteration each lane
// with a ray takes exactly one free-flight step, and the cheap transitions
// (next channel, next macrocell, next sample) happen in the same iteration.
// The expensive parts (Philox, log, division, the trilinear fetch) thus run
// with nearly all lanes. Because every step owns a fixed pair of words of its
// ray's Philox stream, a lane needs a new Philox block only every second step,
// and lanes that start together stay in phase.
//
// Two schedules use this loop:
//   refill = false: a workgroup owns one 16x16 pixel block (waves own 8x8
//     quads), like the one-thread-per-pixel kernel;
//   refill = true: a resident-sized grid whose waves refill idle lanes from
//     eight pixel queues (one per XCD, each a contiguous run of the Hilbert
//     block order; a wave steals from the others when its own runs dry).
// Per pixel the operations and random words are exactly those of the
// one-thread-per-pixel kernel, so all schedules give identical images.
#include "mdwt_internal.hpp"
#include "render_common.hpp"

namespace mdwt {
namespace {

enum : int { S_IDLE = 0, S_STEP = 1, S_CELL = 2, S_START = 3, S_FINISH = 4, S_PASS = 5 };

struct Lane {
  int px, py, s;
  uint32_t pixel;
  float ar, ag, ab;  // pixel sums over samples
  uint32_t k;        // next free-flight step since the last stream seek
  uint32_t base;     // stream block offset (-m 0 / -m 2: visit * 16 + channel)
  uint32_t nv;       // macrocells entered so far (-m 0: in the channel's pass)
  uint32_t b0, b1, b2, b3;
  V3 wdir;
  float w2g, g2w, t1w;
  Dda d;
  float ct0, ct1, minT, t, maj, tmh;
  int cid, c;
  uint32_t mask;
  float m0, m1, m2, m3;  // majorants of the current cell (at most 4 channels)
  bool hit;
  int hc;  // channel of the hit
  float hr, hg, hb;
};

constexpr int kRefill = 16;  // refill a wave once this many lanes are idle

template <int DT>
__device__ __forceinline__ float cell_maj(const Ctx<DT>& X, const Lane& L, int c) {
  if (X.C <= 4) return c == 0 ? L.m0 : (c == 1 ? L.m1 : (c == 2 ? L.m2 : L.m3));
  return X.maj[size_t(L.cid) * X.C + c];
}

template <int MODE, int DT, bool COUNT, bool REFILL>
__global__ __launch_bounds__(256) void render_flat(RenderParams P) {
  extern __shared__ float4 s_tf[];
  __shared__ ChanTf s_ch[kMaxChannels];
  load_tf_lds(P, s_tf, s_ch);
  __syncthreads();
  // clear the other queue set for the next launch (it starts after this one
  // has finished, on the same stream)
  if (REFILL && blockIdx.x == 0 && threadIdx.x < 8)
    P.fr.queues[((P.fr.launch_seq + 1u) & 1u) * 8u + threadIdx.x] = 0u;

  Ctx<DT> X;
  X.init(P, s_tf, s_ch);
  const int C = X.C;
  const V3 corg = load3(P.cam.org), blo = load3(P.vol.blo), bhi = load3(P.vol.bhi);
  const V3 ext = bhi - blo;
  const V3 w2u = v3(1.f / ext.x, 1.f / ext.y, 1.f / ext.z);
  const uint32_t seed = P.fr.seed, frame = P.fr.frame_id;
  const int W = P.fr.width, H = P.fr.height, spp = P.fr.spp;
  const int lane = threadIdx.x & 63;
  const uint32_t nb = uint32_t(P.fr.n_blocks);
  unsigned* queue = P.fr.queues + (P.fr.launch_seq & 1u) * 8u;
  const uint32_t home = blockIdx.x & 7u;

  Counts cnt = {0, 0, 0, 0};
  Lane L;
  L.s = 0;
  int st = S_IDLE;
  bool drained = REFILL ? false : true;
  if (!REFILL) {
    const uint32_t bxy = P.fr.block_order[block_slot(blockIdx.x, nb)];
    const int wave = threadIdx.x >> 6;
    L.px = int(bxy & 0xffffu) * kBlockPixels + (wave & 1) * 8 + (lane & 7);
    L.py = int(bxy >> 16) * kBlockPixels + (wave >> 1) * 8 + (lane >> 3);
    if (L.px < W && L.py < H) {
      L.pixel = uint32_t(L.px) + uint32_t(W) * uint32_t(L.py);
      L.ar = L.ag = L.ab = 0.f;
      st = S_START;
    }
  }

  while (true) {
    // ---------------------------------------------------------- refill --
    if (REFILL) {
      const uint64_t idle = __ballot(st == S_IDLE);
      const int n_idle = __popcll(idle);
      if (drained && n_idle == 64) break;
      if (!drained && n_idle >= kRefill) {
        // queue qi holds the pixels of blocks qi, qi + 8, qi + 16, ... of the
        // Hilbert block order (so every queue gets an even share of heavy
        // regions); v -> block (v >> 8) * 8 + qi, pixel v & 255 in the block
        uint32_t base = 0, avail = 0, qsel = 0;
        if (lane == 0) {
          for (uint32_t q = 0; q < 8u; ++q) {
            const uint32_t qi = (home + q) & 7u;
            const uint32_t size = ((nb + 7u - qi) / 8u) * 256u;
            if (size == 0u) continue;
            const uint32_t old = atomicAdd(&queue[qi], uint32_t(n_idle));
            if (old < size) {
              base = old;
              avail = min(uint32_t(n_idle), size - old);
              qsel = qi;
              break;
            }
          }
        }
        base = __shfl(base, 0, 64);
        avail = __shfl(avail, 0, 64);
        qsel = __shfl(qsel, 0, 64);
        if (avail == 0) drained = true;
        if (st == S_IDLE) {
          const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32),
                                                          __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
          if (rank < avail) {
            const uint32_t v = base + rank;
            const uint32_t i = (((v >> 8) * 8u + qsel) << 8) | (v & 255u);
            const uint32_t bxy = P.fr.block_order[i >> 8];
            const uint32_t q = (i >> 6) & 3u, l = i & 63u;
            L.px = int(bxy & 0xffffu) * kBlockPixels + int(q & 1u) * 8 + int(l & 7u);
            L.py = int(bxy >> 16) * kBlockPixels + int(q >> 1) * 8 + int(l >> 3);
            if (L.px < W && L.py < H) {
              L.pixel = uint32_t(L.px) + uint32_t(W) * uint32_t(L.py);
              L.s = 0;
              L.ar = L.ag = L.ab = 0.f;
              st = S_START;
            }
          }
        }
      }
    } else {
      if (__ballot(st != S_IDLE) == 0) break;
    }

    // ------------------------------------------- one free-flight step --
    if (st == S_STEP) {
      const uint32_t cur = L.k;
      L.k = cur + 1u;
      if ((cur & 1u) == 0u) {
        uint32_t o[4];
        philox(frame, uint32_t(L.s), 1u + L.base, (cur >> 1) << 1, seed, L.pixel, o);
        L.b0 = o[0];
        L.b1 = o[1];
        L.b2 = o[2];
        L.b3 = o[3];
      }
      const float u1 = to_unit((cur & 1u) ? L.b2 : L.b0);
      const float lg = plogf_poly(1.0f - u1);
      L.t = L.t - (lg / L.maj) * X.dt * L.w2g;
      if (MODE == 2) {  // altMultiMajDTCH (deviceCode.cu:800-848)
        bool next_ch = false;
        if (L.t >= L.minT) {
          next_ch = true;
        } else {
          const float tWorld = L.t * L.g2w;
          const V3 xT = (corg + tWorld * L.wdir) * w2u;
          if (tWorld >= L.t1w) {
            cnt.rejections++;
            if (L.c == C - 1)
              st = S_FINISH;
            else
              next_ch = true;
          } else {
            const float thr = to_unit((cur & 1u) ? L.b3 : L.b1) * L.maj;
            const float value = X.sample(xT, L.c);
            cnt.samples++;
            if (!(value != value)) {
              float r2;
              const float a = X.tf_alpha(value, L.c, r2);
              if (a > 0.0f && thr < a) {
                HitRec h;
                X.tf_rgb(r2, L.c, h);
                L.hit = true;
                L.hc = L.c;
                L.hr = h.r;
                L.hg = h.g;
                L.hb = h.b;
                L.minT = fminf(L.t, L.minT);
              }
            }
            cnt.rejections++;
          }
        }
        if (next_ch) {
          L.mask &= ~((2u << L.c) - 1u);
          if (L.mask) {
            L.c = __builtin_ctz(L.mask);
            L.maj = cell_maj(X, L, L.c);
            L.t = L.ct0;
            L.base = ((L.nv - 1u) << 4) | uint32_t(L.c);
            L.k = 0;
          } else if (L.hit) {
            st = S_FINISH;
          } else {
            st = L.d.advance() ? S_CELL : S_FINISH;
          }
        }
      } else if (MODE == 0) {  // baseLineDTCH (deviceCode.cu:906-963)
        if (L.t >= L.ct1) {
          st = L.d.advance() ? S_CELL : S_PASS;
        } else if (L.tmh < L.t) {
          st = S_PASS;
        } else {
          const float tWorld = L.t * L.g2w;
          const V3 xT = (corg + tWorld * L.wdir) * w2u;
          if (tWorld >= L.t1w) {
            cnt.rejections++;
            st = S_PASS;
          } else {
            const float thr = to_unit((cur & 1u) ? L.b3 : L.b1) * L.maj;
            const float value = X.sample(xT, L.c);
            cnt.samples++;
            if (!(value != value)) {
              float r2;
              const float a = X.tf_alpha(value, L.c, r2);
              if (a > 0.0f && thr < a) {
                HitRec h;
                X.tf_rgb(r2, L.c, h);
                L.hit = true;
                L.hc = L.c;
                L.hr = h.r;
                L.hg = h.g;
                L.hb = h.b;
                L.tmh = fminf(L.tmh, L.t);
                st = S_PASS;
              } else {
                cnt.rejections++;
              }
            }
          }
        }
      } else {  // cummilativeDTCH (deviceCode.cu:556-612)
        if (L.t >= L.ct1) {
          st = L.d.advance() ? S_CELL : S_FINISH;
        } else {
          const float tWorld = L.t * L.g2w;
          const V3 xT = (corg + tWorld * L.wdir) * w2u;
          if (tWorld >= L.t1w) {
            cnt.rejections++;
            st = S_FINISH;
          } else {
            float thr = to_unit((cur & 1u) ? L.b3 : L.b1) * L.maj;
            bool hit = false;
            for (int c = 0; c < C; ++c) {
              const float value = X.sample(xT, c);
              cnt.samples++;
              if (value != value) continue;
              float r2;
              const float a = X.tf_alpha(value, c, r2);
              if (a > 0.0f && thr < a) {
                HitRec h;
                X.tf_rgb(r2, c, h);
                L.hit = true;
                L.hc = c;
                L.hr = h.r;
                L.hg = h.g;
                L.hb = h.b;
                hit = true;
                break;
              }
              thr -= a;
            }
            if (hit)
              st = S_FINISH;
            else
              cnt.rejections++;
          }
        }
      }
    }

    // ------------------------------- -m 0: the next channel's DDA pass --
    if (MODE == 0 && st == S_PASS) {
      L.c += 1;
      L.nv = 0;
      if (L.c >= C) {
        st = S_FINISH;
      } else {
        const GridRay g = make_grid_ray(corg, L.wdir, blo, bhi, X.mx, X.my, X.mz);
        st = L.d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz) ? S_CELL : S_FINISH;
      }
    }

    // ------------------------ finish a sample / start the next sample --
    while (st == S_FINISH || st == S_START) {
      if (st == S_FINISH) {
        // a Woodcock hit is opaque, so each sample adds exactly 1 to alpha
        float fa_unused = 0.f;
        composite_sample(P, L.hit, L.hr, L.hg, L.hb, 1.f, 1.f, 0.f, 1.f, L.hc, L.ar, L.ag, L.ab,
                         fa_unused);
        L.s += 1;
        if (L.s < spp) {
          st = S_START;
        } else {
          float fa = 1.f;
          for (int i = 0; i < spp; ++i) fa = fa + 1.0f;
          if (!COUNT)
            finish_pixel(P, P.fr.accum, P.fr.m2, P.fr.accum_id, spp, L.pixel, L.ar, L.ag, L.ab, fa);
          st = S_IDLE;
        }
      } else {
        L.k = 0;
        L.base = 0;
        L.nv = 0;
        L.hit = false;
        L.hc = 0;
        cnt.rays++;
        uint32_t o[4];
        philox(frame, uint32_t(L.s), 0u, 0u, seed, L.pixel, o);
        V3 org, dir;
        camera_ray(P.cam, L.px, L.py, W, H, to_unit(o[0]), org, dir);
        float t0, t1;
        st = S_FINISH;
        if (clip_root(org, dir, blo, bhi, t0, t1)) {
          L.wdir = dir;
          L.t1w = t1;
          const GridRay g = make_grid_ray(org, dir, blo, bhi, X.mx, X.my, X.mz);
          L.w2g = g.worldToGridT;
          L.g2w = g.gridToWorldT;
          if (L.d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz)) {
            L.c = 0;
            L.tmh = 1e20f;
            st = S_CELL;
          }
        }
      }
    }

    // -------------------- enter macrocells until one has work (dda.h) --
    while (st == S_CELL) {
      cnt.visits++;
      L.d.range(L.ct0, L.ct1);
      L.cid = L.d.cx + L.d.cy * X.mx + L.d.cz * X.mx * X.my;
      if (MODE == 2) {
        uint32_t m = 0;
        if (C <= 4) {
          const float* mp = X.maj + size_t(L.cid) * C;
          if (C == 4) {
            const float4 v = *reinterpret_cast<const float4*>(mp);
            L.m0 = v.x;
            L.m1 = v.y;
            L.m2 = v.z;
            L.m3 = v.w;
          } else {
            L.m0 = mp[0];
            L.m1 = C > 1 ? mp[1] : 0.f;
            L.m2 = C > 2 ? mp[2] : 0.f;
            L.m3 = 0.f;
          }
          m = (L.m0 != 0.0f ? 1u : 0u) | (L.m1 != 0.0f ? 2u : 0u) | (L.m2 != 0.0f ? 4u : 0u) |
              (L.m3 != 0.0f ? 8u : 0u);
        } else {
          const float* mp = X.maj + size_t(L.cid) * C;
          for (int c = 0; c < C; ++c) m |= (mp[c] == 0.0f ? 0u : 1u) << c;
        }
        L.nv += 1;  // this macrocell is visit nv - 1
        if (m) {
          L.mask = m;
          L.c = __builtin_ctz(m);
          L.maj = cell_maj(X, L, L.c);
          L.t = L.ct0;
          L.minT = L.ct1;
          L.base = ((L.nv - 1u) << 4) | uint32_t(L.c);
          L.k = 0;
          st = S_STEP;
        } else if (!L.d.advance()) {
          st = S_FINISH;
        }
      } else if (MODE == 0) {
        L.maj = X.maj[size_t(L.cid) * C + L.c];
        L.nv += 1;  // this macrocell is visit nv - 1 of the channel's pass
        if (L.maj != 0.00f) {
          L.t = L.ct0;
          L.base = ((L.nv - 1u) << 4) | uint32_t(L.c);
          L.k = 0;
          st = S_STEP;
        } else if (!L.d.advance()) {
          st = S_PASS;
        }
      } else {
        L.maj = X.maj[L.cid];
        if (L.maj != 0.00f) {
          L.t = L.ct0;
          st = S_STEP;
        } else if (!L.d.advance()) {
          st = S_FINISH;
        }
      }
    }
  }
  if (COUNT) flush_counts(P, cnt);
}

template <int MODE, int DT, bool COUNT>
RenderFn flat_fn(bool refill) {
  return refill ? render_flat<MODE, DT, COUNT, true> : render_flat<MODE, DT, COUNT, false>;
}

template <int MODE, bool COUNT>
RenderFn flat_fn_dt(int dtype, bool refill) {
  switch (dtype) {
    case MDWT_DT_U8: return flat_fn<MODE, MDWT_DT_U8, COUNT>(refill);
    case MDWT_DT_U16: return flat_fn<MODE, MDWT_DT_U16, COUNT>(refill);
    case MDWT_DT_F16: return flat_fn<MODE, MDWT_DT_F16, COUNT>(refill);
    default: return flat_fn<MODE, MDWT_DT_F32, COUNT>(refill);
  }
}

}  // namespace

RenderFn select_flat_kernel(int mode, int dtype, bool count, bool refill) {
  if (mode == 0) return count ? flat_fn_dt<0, true>(dtype, refill) : flat_fn_dt<0, false>(dtype, refill);
  if (mode == 2) return count ? flat_fn_dt<2, true>(dtype, refill) : flat_fn_dt<2, false>(dtype, refill);
  return count ? flat_fn_dt<3, true>(dtype, refill) : flat_fn_dt<3, false>(dtype, refill);
}

}  // namespace mdwt
