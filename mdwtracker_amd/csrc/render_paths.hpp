// Device code shared by the frame kernels of kernels_render.hip (one view)
// and kernels_batch.hip (batched views): the Woodcock tracking loops of every
// mode, blend and ray-marcher paths, shading, and the per-block renderer.
//
// Reference: mainRG (deviceCode.cu:283-479), the closest-hit programs
// baseLineDTCH (858-970), multiMajDTCH (619-745), altMultiMajDTCH (747-856),
// cummilativeDTCH (513-617), blendDTCH (972-1065), rayMarcherCH (1067-1157)
// and compRayMarcherCH (1159-1276), running dda::dda3 (dda.h:82-203).
#pragma once
#include "render_common.hpp"

namespace mdwt {
namespace {

// Majorants of up to four channels of one macrocell: one 16-byte load for four
// channels (the grid is cell-major, 16-byte aligned per cell when C = 4).
__device__ __forceinline__ float4 load_cell_maj(const float* maj, int cid, int C) {
  if (C == 4) return *reinterpret_cast<const float4*>(maj + size_t(cid) * 4);
  const float* p = maj + size_t(cid) * C;
  return make_float4(p[0], C > 1 ? p[1] : 0.f, C > 2 ? p[2] : 0.f, 0.f);
}
__device__ __forceinline__ float pick(const float4& m, int c) {
  return c == 0 ? m.x : (c == 1 ? m.y : (c == 2 ? m.z : m.w));
}

// -m 2, altMultiMajDTCH (deviceCode.cu:747-856). With C <= 4 the next cell's
// majorants are requested before the current cell's channels are tracked, so
// the load latency overlaps the free-flight steps instead of preceding them.
template <int DT>
__device__ void trace_sync(const Ctx<DT>& X, const GridRay& g, float prd_t1, Rng& rng, HitRec& h,
                           Counts& cnt) {
  Dda d;
  if (!d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz)) return;
  const float step = X.dt;
  const bool regs = X.C <= 4;
  int cid = d.cx + d.cy * X.mx + d.cz * X.mx * X.my;
  float4 mcur = regs ? load_cell_maj(X.maj, cid, X.C) : make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t visit = 0;  // macrocells entered so far: the stream's block offset
  while (true) {
    cnt.visits++;
    float t0, t1;
    d.range(t0, t1);
    Dda dn = d;
    const bool more = dn.advance();
    const int ncid = dn.cx + dn.cy * X.mx + dn.cz * X.mx * X.my;
    const float4 mnext = (regs && more) ? load_cell_maj(X.maj, ncid, X.C) : mcur;
    float minT = t1;
    bool terminate = false;
    // first distances of up to four channels from one cell block (all lanes
    // of the wave compute it together, once per cell and channel group)
    uint32_t cw[4];
    if (regs) rng.cell_block(visit, 0u, cw);
    for (int c = 0; c < X.C; ++c) {
      const float majorant = regs ? pick(mcur, c) : X.maj[size_t(cid) * X.C + c];
      float t = t0;
      if (!regs && (c & 3) == 0) rng.cell_block(visit, uint32_t(c >> 2), cw);
      if (majorant == 0.0f) continue;
      float u = to_unit(pick_word(cw, uint32_t(c & 3)));
      rng.start_pairs(visit, uint32_t(c));
      while (true) {
        t = t - (plogf_poly(1.0f - u) / majorant) * step * g.worldToGridT;
        if (t >= minT) break;
        const float tWorld = t * g.gridToWorldT;
        const V3 xT = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
        if (tWorld >= prd_t1) {
          cnt.rejections++;
          if (c == X.C - 1) terminate = true;
          break;
        }
        const float thr = rng.pair_thr() * majorant;
        u = rng.pair_dist();
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
    if (!more) return;
    d = dn;
    cid = ncid;
    mcur = mnext;
    ++visit;
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
    float majorant = X.maj[size_t(d.cx + d.cy * X.mx + d.cz * X.mx * X.my) * X.C + c];
    uint32_t visit = 0;
    while (true) {
      cnt.visits++;
      float t0, t1;
      d.range(t0, t1);
      // next cell's majorant requested before this cell is tracked
      Dda dn = d;
      const bool more = dn.advance();
      const float mnext =
          more ? X.maj[size_t(dn.cx + dn.cy * X.mx + dn.cz * X.mx * X.my) * X.C + c] : 0.f;
      bool stop = false;
      if (majorant != 0.00f) {
        uint32_t cw[4];
        rng.cell_block(visit, uint32_t(c >> 2), cw);
        float u = to_unit(pick_word(cw, uint32_t(c & 3)));
        rng.start_pairs(visit, uint32_t(c));
        float t = t0;
        while (true) {
          t = t - (plogf_poly(1.0f - u) / majorant) * step * g.worldToGridT;
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
          const float thr = rng.pair_thr() * majorant;
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
              tMaxHit = fminf(tMaxHit, t);
              stop = true;
              break;
            }
            cnt.rejections++;
          }
          u = rng.pair_dist();
        }
      }
      if (stop) break;
      if (!more) break;
      d = dn;
      majorant = mnext;
      ++visit;
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
  float majorant = X.maj[d.cx + d.cy * X.mx + d.cz * X.mx * X.my];
  uint32_t visit = 0u;  // the (visit, channel 0) positions of the stream
  while (true) {
    cnt.visits++;
    float t0, t1;
    d.range(t0, t1);
    Dda dn = d;
    const bool more = dn.advance();
    const float mnext = more ? X.maj[dn.cx + dn.cy * X.mx + dn.cz * X.mx * X.my] : 0.f;
    if (majorant != 0.00f) {
      float t = t0;
      uint32_t cw[4];
      rng.cell_block(visit, 0u, cw);
      rng.start_pairs(visit, 0u);
      float u = to_unit(cw[0]);
      while (true) {
        t = t - (plogf_poly(1.0f - u) / majorant) * step * g.worldToGridT;
        if (t >= t1) break;
        const float tWorld = t * g.gridToWorldT;
        const V3 xT = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
        if (tWorld >= prd_t1) {
          cnt.rejections++;
          return;
        }
        float thr = rng.pair_thr() * majorant;
        u = rng.pair_dist();
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
    if (!more) return;
    d = dn;
    majorant = mnext;
    ++visit;
  }
}

// -m 1, multiMajDTCH (deviceCode.cu:619-745): per cell one exponential clock
// per channel, the earliest fires; only the fired channel's clock advances.
// Channel c's clock starts at the first distance of (visit, c) and its j-th
// event takes pair j of the (visit, c) stream. Clocks, majorants and pair
// counts live in registers: the loops run over kMaxChannels with constant
// indices and stop at X.C, so no private-memory arrays appear.
template <int DT>
__device__ void trace_multi(const Ctx<DT>& X, const GridRay& g, float prd_t1, Rng& rng, HitRec& h,
                            Counts& cnt) {
  Dda d;
  if (!d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz)) return;
  const float step = X.dt;
  float maj[kMaxChannels], ts[kMaxChannels];
  uint32_t pj[kMaxChannels];
  uint32_t visit = 0u;
  while (true) {
    cnt.visits++;
    float t0, t1;
    d.range(t0, t1);
    const int cid = d.cx + d.cy * X.mx + d.cz * X.mx * X.my;
    float sum = 0.0f;
    if (X.C <= 4) {
      // one 16-byte load for the cell's majorants instead of one dependent
      // load per channel; the sum keeps the channel order
      const float4 m4 = load_cell_maj(X.maj, cid, X.C);
#pragma unroll
      for (int i = 0; i < kMaxChannels; ++i) {
        maj[i] = i < 4 && i < X.C ? pick(m4, i) : 0.0f;
        if (i < X.C) sum += maj[i];
        ts[i] = t0;
        pj[i] = 0u;
      }
    } else {
#pragma unroll
      for (int i = 0; i < kMaxChannels; ++i) {
        if (i < X.C) {
          maj[i] = X.maj[size_t(cid) * X.C + i];
          sum += maj[i];
        } else {
          maj[i] = 0.0f;
        }
        ts[i] = t0;
        pj[i] = 0u;
      }
    }
    if (!(sum <= 0.0000001f)) {
      uint32_t cw[4] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int i = 0; i < kMaxChannels; ++i) {
        if (i < X.C) {
          if ((i & 3) == 0) rng.cell_block(visit, uint32_t(i >> 2), cw);
          if (maj[i] > 0.0f)
            ts[i] = ts[i] - (plogf_poly(1.0f - to_unit(cw[i & 3])) / maj[i]) * step *
                                g.worldToGridT;
          else
            ts[i] = 1e20f;
        }
      }
      while (true) {
        float minT = ts[0], mj = maj[0];
        int sel = 0;
        uint32_t j = pj[0];
#pragma unroll
        for (int i = 1; i < kMaxChannels; ++i) {
          if (i < X.C && ts[i] < minT && maj[i] > 0.f) {
            sel = i;
            minT = ts[i];
            mj = maj[i];
            j = pj[i];
          }
        }
        if (minT >= t1) break;
        const float tWorld = minT * g.gridToWorldT;
        const V3 xT = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
        if (tWorld >= prd_t1) {
          cnt.rejections++;
          return;
        }
        float th, rand;
        rng.vc_pair(visit, uint32_t(sel), j, th, rand);
        const float thr = th * mj;
        const float value = X.sample(xT, sel);
        cnt.samples++;
        if (!(value != value)) {
          float r2;
          const float a = X.tf_alpha(value, sel, r2);
          if (a > 0.0f && thr < a) {
            h.missed = false;
            h.tHit = tWorld;
            h.channel = sel;
            X.tf_rgb(r2, sel, h);
            return;
          }
        }
        cnt.rejections++;
        const float nt = minT - (plogf_poly(1.0f - rand) / mj) * step * g.worldToGridT;
#pragma unroll
        for (int i = 0; i < kMaxChannels; ++i) {
          if (i == sel) {
            ts[i] = nt;
            pj[i] = j + 1u;
          }
        }
      }
    }
    if (!d.advance()) return;
    ++visit;
  }
}

// blendChannels (deviceCode.cu:222-241): maxOpacityBlend (165-179) for -m 4
// and 7, mixBlend (182-201) for -m 5 and 8, densityBlend (204-220) for -m 6;
// (0,0,0,0) outside the unit cube.
// The ray marchers (-m 6-8) sample every channel at every step: with at most
// four channels all channels are sampled first, so their voxel loads are in
// flight together instead of one channel's after another's; the blend then
// runs in channel order on the same values. Blend tracking (-m 4/5) blends
// only at tentative collisions and keeps the plain loop (fewer registers).
template <int MODE, int DT>
__device__ float4 blend_channels(const Ctx<DT>& X, V3 p) {
  if (p.x < 0.0f || p.x > 1.0f || p.y < 0.0f || p.y > 1.0f || p.z < 0.0f || p.z > 1.0f)
    return make_float4(0.f, 0.f, 0.f, 0.f);
  const bool pre = MODE >= 6 && X.C <= 4;
  float4 v4 = make_float4(0.f, 0.f, 0.f, 0.f);
  if (pre) {
    v4.x = X.sample(p, 0);
    if (X.C > 1) v4.y = X.sample(p, 1);
    if (X.C > 2) v4.z = X.sample(p, 2);
    if (X.C > 3) v4.w = X.sample(p, 3);
  }
  auto value = [&](int c) { return pre ? pick(v4, c) : X.sample(p, c); };
  if (MODE == 6) {
    float sum = 0.0f;
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int c = 0; c < X.C; ++c) {
      const float4 s = X.tf_rgba(value(c), c);
      col.x = col.x + s.w * s.x;
      col.y = col.y + s.w * s.y;
      col.z = col.z + s.w * s.z;
      col.w = col.w + s.w * s.w;
      sum += s.w;
    }
    if (sum <= 0.0f) return make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4(col.x / sum, col.y / sum, col.z / sum, col.w / sum);
  }
  if (MODE == 4 || MODE == 7) {
    float4 best = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int c = 0; c < X.C; ++c) {
      const float4 s = X.tf_rgba(value(c), c);
      if (s.w > best.w) best = s;
    }
    return best;
  }
  float cr = 0.f, cg = 0.f, cb = 0.f, mo = 0.0f;  // -m 5 and 8
  for (int c = 0; c < X.C; ++c) {
    const float4 s = X.tf_rgba(value(c), c);
    cr = cr + s.x * s.w;
    cg = cg + s.y * s.w;
    cb = cb + s.z * s.w;
    if (s.w > mo) mo = s.w;
  }
  if (mo <= 0.0f) return make_float4(0.f, 0.f, 0.f, 0.f);
  return make_float4(cr / mo, cg / mo, cb / mo, mo);
}

// -m 4 / 5, blendDTCH (deviceCode.cu:972-1065): Woodcock tracking on the
// summed majorant with a blended sample; the hit keeps channel 0 (the
// reference never sets the payload's channel in this program).
template <int MODE, int DT>
__device__ void trace_blend(const Ctx<DT>& X, const GridRay& g, float prd_t1, Rng& rng, HitRec& h,
                            Counts& cnt) {
  Dda d;
  if (!d.begin(g.org, g.dir, 1e20f, X.mx, X.my, X.mz)) return;
  const float step = X.dt;
  float majorant = X.maj[d.cx + d.cy * X.mx + d.cz * X.mx * X.my];
  uint32_t visit = 0u;  // the (visit, channel 0) positions of the stream, as -m 3
  while (true) {
    cnt.visits++;
    float t0, t1;
    d.range(t0, t1);
    Dda dn = d;
    const bool more = dn.advance();
    const float mnext = more ? X.maj[dn.cx + dn.cy * X.mx + dn.cz * X.mx * X.my] : 0.f;
    if (majorant != 0.00f) {
      float t = t0;
      uint32_t cw[4];
      rng.cell_block(visit, 0u, cw);
      rng.start_pairs(visit, 0u);
      float u = to_unit(cw[0]);
      while (true) {
        t = t - (plogf_poly(1.0f - u) / majorant) * step * g.worldToGridT;
        if (t >= t1) break;
        const float tWorld = t * g.gridToWorldT;
        const V3 xT = (g.worldOrg + tWorld * g.worldDir) * g.worldToUnit;
        if (tWorld >= prd_t1) {
          cnt.rejections++;
          return;
        }
        const float4 s = blend_channels<MODE, DT>(X, xT);
        cnt.samples++;
        const float th = rng.pair_thr();
        u = rng.pair_dist();
        if (s.w > 0.0f && th * majorant < s.w) {
          h.missed = false;
          h.tHit = tWorld;
          h.channel = 0;
          h.r = s.x;
          h.g = s.y;
          h.b = s.z;
          return;
        }
        cnt.rejections++;
      }
    }
    if (!more) return;
    d = dn;
    majorant = mnext;
    ++visit;
  }
}

// -m 6 / 7 / 8, rayMarcherCH (deviceCode.cu:1067-1157) with shadows off:
// front-to-back compositing of blended samples every dt from the root-box
// entry; the sample position trails t by one step and starts at the ray
// origin (the reference moves pos after sampling).
template <int MODE, int DT, bool LIT>
__device__ void march_blend(const RenderParams& P, const Ctx<DT>& X, V3 org, V3 dir0, V3 w2u,
                            float prd_t0, float prd_t1, HitRec& h, Counts& cnt) {
  const float dt = X.dt;
  const V3 dir = normalize3(dir0);
  V3 pos = org;
  float alpha = 0.0f;
  float cr = 0.f, cg = 0.f, cb = 0.f;
  const bool shadows = LIT && P.fr.sh.shadows != 0;
  float hiC = 0.0f, hiT = 0.0f;
  for (float t = prd_t0; (t < prd_t1) && (alpha < 0.99f); t += dt) {
    const V3 posTex = pos * w2u;
    const float4 b = blend_channels<MODE, DT>(X, posTex);
    cnt.samples += uint32_t(X.C);
    const float k = 1.f - alpha;
    cr = cr + b.x * b.w * k;
    cg = cg + b.y * b.w * k;
    cb = cb + b.z * b.w * k;
    alpha = alpha + (1.f - alpha) * b.w;
    if (shadows && hiC < b.w) {
      hiC = b.w;
      hiT = t;
    }
    pos = org + dir * t;
  }
  // deviceCode.cu:1117-1154: one marched shadow ray from the most opaque
  // sample towards the light; its first sample sits one world unit from that
  // point (posSh = orgSh + dirSh), as in the reference.
  float sh = 1.0f;
  if (shadows && hiC > 0.0f) {
    const V3 orgSh = org + dir * hiT;
    const V3 dirSh = v3(-P.fr.sh.light_dir[0], -P.fr.sh.light_dir[1], -P.fr.sh.light_dir[2]);
    float alphaSh = 0.f;
    V3 posSh = orgSh + dirSh;
    const V3 df = v3(1.f / dirSh.x, 1.f / dirSh.y, 1.f / dirSh.z);
    const float t1 = (P.vol.blo[0] - orgSh.x) * df.x, t2 = (P.vol.bhi[0] - orgSh.x) * df.x;
    const float t3 = (P.vol.blo[1] - orgSh.y) * df.y, t4 = (P.vol.bhi[1] - orgSh.y) * df.y;
    const float t5 = (P.vol.blo[2] - orgSh.z) * df.z, t6 = (P.vol.bhi[2] - orgSh.z) * df.z;
    float thit0 = fmaxf(fmaxf(fminf(t1, t2), fminf(t3, t4)), fminf(t5, t6));
    const float thit1 = fminf(fminf(fmaxf(t1, t2), fmaxf(t3, t4)), fmaxf(t5, t6));
    if ((thit1 >= 0) && (thit0 < thit1)) {
      thit0 = fmaxf(thit0, 0.00001f);
      if (sqrtf(dot3(dirSh, dirSh)) < 1e-5f) thit0 = thit1;
      for (float tSh = thit0; (tSh < thit1) && (alphaSh < 0.99f); tSh += dt) {
        const V3 posTexSh = posSh * w2u;
        const float opacitySh = blend_channels<MODE, DT>(X, posTexSh).w;
        alphaSh = alphaSh + (1.f - alphaSh) * opacitySh;
        posSh = orgSh + dirSh * tSh;
      }
      cnt.samples += uint32_t(X.C);
    }
    sh = (1.f - P.fr.ambient) * (1.f - alphaSh) + P.fr.ambient;
  }
  h.missed = false;
  h.r = cr * sh;
  h.g = cg * sh;
  h.b = cb * sh;
  h.a = alpha;
}

// -m 9, compRayMarcherCH (deviceCode.cu:1159-1276): each channel marched and
// composited on its own, then mixed by opacity. The sample position is not
// reset between channels (channel n > 0 starts where channel n-1 stopped), as
// in the reference.
template <int DT>
__device__ void march_comp(const Ctx<DT>& X, V3 org, V3 dir0, V3 w2u, float prd_t0, float prd_t1,
                           HitRec& h, Counts& cnt) {
  const float dt = X.dt;
  const V3 dir = normalize3(dir0);
  V3 pos = org;
  float asum = 0.0f, fr = 0.f, fg = 0.f, fb = 0.f;
  float al[kMaxChannels], rr[kMaxChannels], gg[kMaxChannels], bb[kMaxChannels];
#pragma unroll
  for (int n = 0; n < kMaxChannels; ++n) {
    float a = 0.0f, r = 0.f, g = 0.f, b = 0.f;
    if (n < X.C) {
      for (float t = prd_t0; (t < prd_t1) && (a < 0.99f); t += dt) {
        const V3 posTex = pos * w2u;
        const float v = X.sample(posTex, n);
        float4 sc = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!(v != v)) {
          sc = X.tf_rgba(v, n);
          cnt.samples++;
        }
        const float k = 1.f - a;
        r = r + sc.x * sc.w * k;
        g = g + sc.y * sc.w * k;
        b = b + sc.z * sc.w * k;
        a = a + (1.f - a) * sc.w;
        pos = org + dir * t;
      }
    }
    al[n] = a;
    rr[n] = r;
    gg[n] = g;
    bb[n] = b;
  }
#pragma unroll
  for (int n = 0; n < kMaxChannels; ++n)
    if (n < X.C) asum += al[n];
  if (asum > 0.0f) {
#pragma unroll
    for (int n = 0; n < kMaxChannels; ++n) {
      if (n < X.C) {
        fr = fr + rr[n] * al[n] / asum;
        fg = fg + gg[n] * al[n] / asum;
        fb = fb + bb[n] * al[n] / asum;
      }
    }
  }
  h.missed = false;
  h.r = fr;
  h.g = fg;
  h.b = fb;
  h.a = asum / float(X.C);
}

// The mode's closest-hit program for the Woodcock modes (0-5).
template <int MODE, int DT>
__device__ __forceinline__ void trace_mode(const Ctx<DT>& X, const GridRay& g, float t1, Rng& rng,
                                           HitRec& h, Counts& cnt) {
  if (MODE == 0)
    trace_ndda<DT>(X, g, t1, rng, h, cnt);
  else if (MODE == 1)
    trace_multi<DT>(X, g, t1, rng, h, cnt);
  else if (MODE == 2)
    trace_sync<DT>(X, g, t1, rng, h, cnt);
  else if (MODE == 3)
    trace_cumulative<DT>(X, g, t1, rng, h, cnt);
  else
    trace_blend<MODE, DT>(X, g, t1, rng, h, cnt);
}

// Gradient (Blinn-Phong) shading of a Woodcock hit, deviceCode.cu:326-381,
// with the hit channel's material. pow is exp(y log x) (ppowf).
template <int DT>
__device__ void shade_gradient(const RenderParams& P, const Ctx<DT>& X, V3 org, V3 dir,
                               const HitRec& h, float& diffuse, float& specular) {
  if (!(sqrtf(h.r * h.r + h.g * h.g + h.b * h.b) > 0.001f)) return;
  const int m = h.channel;
  const V3 blo = load3(P.vol.blo), bhi = load3(P.vol.bhi);
  const V3 w2u = v3(1.f / (bhi.x - blo.x), 1.f / (bhi.y - blo.y), 1.f / (bhi.z - blo.z));
  const V3 tex = (org + h.tHit * dir) * w2u;
  const V3 grad = X.gradient(tex, P.fr.sh.grad_step[m], m);
  if (sqrtf(dot3(grad, grad)) > 1e-6f) {
    const V3 N = normalize3(v3(-grad.x, -grad.y, -grad.z));
    const V3 L = normalize3(load3(P.fr.sh.light_dir));
    const V3 V = normalize3(v3(-dir.x, -dir.y, -dir.z));
    const V3 Hh = normalize3(L + V);
    const float NdotL = fmaxf(dot3(N, L), 0.0f);
    const float NdotH = fmaxf(dot3(N, Hh), 0.0f);
    diffuse = P.fr.sh.diffuse[m] * NdotL;
    if (NdotL > 0.0f && NdotH > 0.001f) {
      const float normalization = (P.fr.sh.shininess[m] + 8.0f) / (8.0f * 3.14159265f);
      specular = P.fr.sh.specular[m] * normalization * ppowf(NdotH, P.fr.sh.shininess[m]);
    } else {
      specular = 0.0f;
    }
  } else {
    diffuse = P.fr.sh.diffuse[m];
    specular = 0.0f;
  }
}

// Denoiser inputs of one sample folded into the pixel's running means:
// first-hit distance (0 on a miss), transmittance (1 on a miss, 0 on a hit)
// and a one-hot of the hit channel, aov_groups(C) float4 per pixel.
__device__ __forceinline__ void aov_sample(const RenderParams& P, const ViewCtx& V, uint32_t pixel,
                                           int s, const HitRec& h) {
  const int ng = aov_groups(P.vol.channels);
  float4* a = reinterpret_cast<float4*>(P.fr.aov) + size_t(ng) * pixel;
  const uint32_t n = (P.fr.accumulate ? V.accum_id * uint32_t(V.spp) : 0u) + uint32_t(s);
  const float k = float(n), k1 = float(n + 1u);
  const bool hit = !h.missed;
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int g = 0; g < ng; ++g) {
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int slot = 4 * g + j;
      v[j] = slot == 0 ? (hit ? h.tHit : 0.f)
                       : slot == 1 ? (hit ? 0.f : 1.f) : (hit && h.channel == slot - 2 ? 1.f : 0.f);
    }
    const float4 o = n ? a[g] : z;
    a[g] = make_float4((v[0] + k * o.x) / k1, (v[1] + k * o.y) / k1, (v[2] + k * o.z) / k1,
                       (v[3] + k * o.w) / k1);
  }
}

// One workgroup's pixel block (bw x bh pixels, one 8x8 quad per wave): every
// pixel's samples, shading and accumulation.
template <int MODE, int DT, bool COUNT, bool LIT>
__device__ __forceinline__ void render_block(const RenderParams& P, const Ctx<DT>& X,
                                             const ViewCtx& V, int bx, int by, Counts& cnt) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int qs = P.fr.bw == 16 ? 1 : 0;  // quads per block row: 1 or 2
  const int px = bx * P.fr.bw + (wave & qs) * 8 + (lane & 7);
  const int py = by * P.fr.bh + (wave >> qs) * 8 + (lane >> 3);
  if (px >= V.W || py >= V.H) return;
  const uint32_t pixel = uint32_t(px) + uint32_t(V.W) * uint32_t(py);
  const V3 blo = load3(P.vol.blo), bhi = load3(P.vol.bhi);
  float fr = 0.f, fg = 0.f, fb = 0.f, fa = 1.f;
  for (int s = 0; s < V.spp; ++s) {
    Rng rng;
    rng.init(V.seed, pixel, V.frame_id, uint32_t(s), 0u);
    float jx, jy;
    rng.jitter2(P.fr.jitter2d != 0, jx, jy);
    V3 org, dir;
    camera_ray(*V.cam, px, py, V.W, V.H, jx, jy, org, dir);
    HitRec h;
    h.missed = true;
    h.a = 1.f;
    h.channel = 0;
    h.tHit = 0.f;
    float diffuse = 1.0f, specular = 0.0f, shadow = 1.0f;
    if (MODE <= 5) {
      // primary ray, then (when shadows are on and it hit) one shadow ray
      // towards the light through the same tracking code, own RNG stream
      // (deviceCode.cu:321, 383-403)
      V3 ro = org, rd = dir;
      for (int pass = 0; pass < (LIT ? 2 : 1); ++pass) {
        HitRec hp;
        hp.missed = true;
        hp.a = 1.f;
        hp.channel = 0;
        float t0, t1;
        if (clip_root(ro, rd, blo, bhi, t0, t1)) {
          const GridRay g = make_grid_ray(ro, rd, blo, bhi, X.mx, X.my, X.mz);
          trace_mode<MODE, DT>(X, g, t1, rng, hp, cnt);
        }
        if (pass == 1) {
          shadow = (1.f - P.fr.ambient) * (1.f - (hp.missed ? 0.0f : 1.0f));
          break;
        }
        h = hp;
        if (!LIT || h.missed || !P.fr.sh.shadows) break;
        ro = org + h.tHit * dir;
        rd = v3(-P.fr.sh.light_dir[0], -P.fr.sh.light_dir[1], -P.fr.sh.light_dir[2]);
        rng.init(V.seed, pixel, V.frame_id, uint32_t(s), 1u);
      }
      if (LIT && !h.missed && P.fr.sh.gradient) shade_gradient(P, X, org, dir, h, diffuse, specular);
    } else {
      float t0, t1;
      if (clip_root(org, dir, blo, bhi, t0, t1)) {
        const V3 w2u = v3(1.f / (bhi.x - blo.x), 1.f / (bhi.y - blo.y), 1.f / (bhi.z - blo.z));
        if (MODE <= 8)
          march_blend<MODE, DT, LIT>(P, X, org, dir, w2u, t0, t1, h, cnt);
        else
          march_comp<DT>(X, org, dir, w2u, t0, t1, h, cnt);
      }
    }
    cnt.rays++;
    composite_sample(P, !h.missed, h.r, h.g, h.b, h.a, diffuse, specular, shadow, h.channel, fr,
                     fg, fb, fa);
    if (LIT && !COUNT && P.fr.aov) aov_sample(P, V, pixel, s, h);
  }
  if (!COUNT) finish_pixel(P, V.accum, V.m2, V.accum_id, V.spp, pixel, fr, fg, fb, fa);
}

}  // namespace
}  // namespace mdwt
