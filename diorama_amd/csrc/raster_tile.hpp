// The per-tile bodies of the rasterizer and its backward pass, shared by the
// single-frame kernels (render.hip) and the multi-view batch kernels (batch.hip).
#pragma once
#include "dio_internal.hpp"

namespace dio {
namespace raster {

// Both rasterizer kernels run one workgroup per 16x16 tile as four independent
// wavefronts, one per 8x8 quadrant (one pixel per lane). A wave stages 64 of the
// tile's splats at a time (one per lane) and parks their blend constants in LDS,
// from where each is read back by the whole wave with a uniform address (a
// broadcast on the LDS pipe; nine v_readlane per splat took VALU issue slots),
// and the forward pass needs no workgroup barrier at all. Each staged splat
// also gets the bounding box of its alpha >= 1/255 ellipse; a wave only visits
// the splats whose box meets its quadrant (a ballot mask), which changes no
// result: outside the box every pixel would skip the splat anyway.

constexpr int kChunk = kWave;
constexpr int kGrads = 9;

// Where a wavefront works: the 8x8 quadrant `quad` (0-3, row-major) of the 16-px
// tile (sx, sy). With a 32-px bin the pair list is the union of its four 16-px
// tiles' lists in the same depth order, and a splat counts for a tile only if its
// 16-px tile rectangle (common.wgsl:37-42) holds the tile: what is left is exactly
// that tile's own list, so every pixel blends the same splats in the same order as
// with 16-px tiles. Only the pixel_last indices (positions in the list) differ.
struct WaveTile {
  uint32_t sx, sy, quad;
};

// One splat of a chunk, staged in a lane's registers. The conic (a, b, c) is
// kept as the coefficients of the Gaussian's exponent in log2 units,
// ha = -a/2 log2(e), hb = -b log2(e), hc = -c/2 log2(e), so G = exp2(power2).
struct Staged {
  float mx, my, ha, hb, hc, op, r, g, b;
  float x0, x1, y0, y1;  // box holding every pixel centre where opacity * G >= 1/255
  uint32_t index;
  uint32_t rect[4];      // the 16-px tile rectangle [x0, x1) x [y0, y1), with kCull
};

constexpr float kLog2e = 1.4426950408889634f;

// The exponent of G at (dx, dy) in log2 units: the forward pass and the backward
// replay both use exactly this, so they agree on every alpha.
__device__ inline float splat_power2(float dx, float dy, float ha, float hb, float hc) {
  return __builtin_fmaf(dx, __builtin_fmaf(ha, dx, hb * dy), (hc * dy) * dy);
}

// Stages pair k's splat; `conic` (optional) receives its raw (a, b, c). With kCull
// it also gets the splat's 16-px tile rectangle in a tiles_x x tiles_y grid.
template <bool kCull = false>
__device__ inline void stage(const float4* __restrict__ splats, const uint32_t* __restrict__ pair_splats, uint32_t k,
                             Staged& s, float4* conic = nullptr, uint32_t tiles_x = 0, uint32_t tiles_y = 0) {
  s.index = pair_splats[k];
  const float4 a = splats[s.index * 3], b = splats[s.index * 3 + 1], c = splats[s.index * 3 + 2];
  if (kCull) tile_rect(a.x, a.y, a.w, tiles_x, tiles_y, s.rect);
  s.mx = a.x; s.my = a.y;
  const float ca = b.x, cb = b.y, cc = b.z;
  if (conic) *conic = make_float4(ca, cb, cc, 0.f);
  s.ha = -0.5f * kLog2e * ca;
  s.hb = -kLog2e * cb;
  s.hc = -0.5f * kLog2e * cc;
  s.r = c.x; s.g = c.y; s.b = c.z; s.op = c.w;
  // opacity * exp(-q) >= 1/255 needs q <= t = ln(255 opacity), with
  // q = (a dx² + c dy²)/2 + b dx dy. That ellipse's half-widths are sqrt(2t Σxx)
  // and sqrt(2t Σyy), Σ = conic⁻¹. Widened by 1% and 0.01 px against rounding.
  const float det = ca * cc - cb * cb;
  const float t = __logf(255.f * s.op) * 1.01f + 0.01f;
  float ex = sqrtf(2.f * t * cc / det) + 0.01f, ey = sqrtf(2.f * t * ca / det) + 0.01f;
  ex = !(det > 0.f) ? INFINITY : (!(t > 0.f) ? -INFINITY : ex);  // degenerate: never cull; too faint: always
  ey = !(det > 0.f) ? INFINITY : (!(t > 0.f) ? -INFINITY : ey);
  s.x0 = s.mx - ex;
  s.x1 = s.mx + ex;
  s.y0 = s.my - ey;
  s.y1 = s.my + ey;
}

__device__ inline bool meets(const Staged& s, float bx0, float bx1, float by0, float by1) {
  return s.x1 >= bx0 && s.x0 <= bx1 && s.y1 >= by0 && s.y0 <= by1;
}

__device__ inline bool holds(const uint32_t rect[4], uint32_t sx, uint32_t sy) {
  return sx >= rect[0] && sx < rect[2] && sy >= rect[1] && sy < rect[3];
}

// A staged splat's blend constants as LDS records: (mean x, mean y, ha, hb),
// (hc, opacity, r, g) and (b, -, -, -). A wave reads one back with a uniform
// address (a broadcast on the LDS pipe) instead of nine v_readlane on the VALU.
// The three arrays share one 16-byte stride, so a record's three reads take one
// address and constant offsets (with b in a 4-byte array, hipcc derived b's
// address from a's with a 64-bit v_mad_u64_u32 per splat). The padding costs the
// forward at 128 or 256 splats per round a quarter of its occupancy (24 KB of
// records per workgroup); at 64 it took the forward from 111.6 to 107.1 µs and the
// backward from 320 to 310.5 µs (1.1M Gaussians, 980x545, one box).
struct SplatRecords {
  float4 a[kChunk], c[kChunk], b[kChunk];
};

__device__ inline void put_record(SplatRecords& r, int slot, const Staged& s) {
  r.a[slot] = make_float4(s.mx, s.my, s.ha, s.hb);
  r.c[slot] = make_float4(s.hc, s.op, s.r, s.g);
  r.b[slot].x = s.b;
}

// Orders a wave's LDS writes before its later reads of other lanes' words.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o, kWave)));
  return v;
}

// rasterize.wgsl:31-100: front-to-back blend; a wave stops once all its pixels
// are nearly opaque.
// One wave's quadrant `wt` of a width x height frame whose pair list is `run`; the
// per-pixel outputs are indexed by the frame's own row-major pixel index. A wave
// stages kPer x 64 pairs per round (kPer per lane) into its own `records`; the
// per-pixel order and every result are the same for any kPer. kCull: `run` is a
// 32-px bin's list (see WaveTile).
// ckpt (16-px tiles, nullable): the segmented backward's checkpoints. Rounds then
// start at absolute multiples of kPer x 64, so every boundary (a multiple of
// 2^seg_shift) starts a round; before it, each pixel's (T, r, g, b) goes to
// ckpt[(b >> seg_shift) * 256 + quadrant * 64 + lane], and the wave returns its
// largest pixel_last (0 without ckpt).
template <int kPer, bool kCull>
__device__ __forceinline__ uint32_t rasterize_wave(const float4* __restrict__ splats,
                                                   const uint32_t* __restrict__ pair_splats, uint2 run, WaveTile wt,
                                                   uint32_t width, uint32_t height, float bg_r, float bg_g,
                                                   float bg_b, uint32_t* __restrict__ target,
                                                   float4* __restrict__ pixel_color,
                                                   uint32_t* __restrict__ pixel_last, SplatRecords (&records)[kPer],
                                                   float4* __restrict__ ckpt = nullptr, uint32_t seg_shift = 0) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t qx = wt.sx * kTile + (wt.quad & 1) * 8, qy = wt.sy * kTile + (wt.quad >> 1) * 8;
  const uint32_t px = qx + (lane & 7), py = qy + (lane >> 3);
  const uint32_t tiles_x = ceil_div(width, kTile), tiles_y = ceil_div(height, kTile);
  const bool inside = px < width && py < height;
  const float cx = static_cast<float>(px) + 0.5f, cy = static_cast<float>(py) + 0.5f;
  const float bx0 = static_cast<float>(qx) + 0.5f, bx1 = bx0 + 7.f;
  const float by0 = static_cast<float>(qy) + 0.5f, by1 = by0 + 7.f;
  // A pixel that has finished keeps its transmittance negated: T > 0 means still
  // blending (T >= 1e-4 then), with no separate flag to carry and test per splat.
  // Pixels outside the frame start finished.
  float T = inside ? 1.f : -1.f, cr = 0.f, cg = 0.f, cb = 0.f;
  uint32_t last = 0;
  constexpr uint32_t kRound = kPer * kChunk;
  const uint32_t first = ckpt && run.x < run.y ? run.x - run.x % kRound : run.x;
  for (uint32_t round = first; round < run.y; round += kRound) {
    if (__all(T < 0.f)) break;
    if (ckpt && round > run.x && (round & ((1u << seg_shift) - 1u)) == 0)
      ckpt[static_cast<size_t>(round >> seg_shift) * (4 * kChunk) + wt.quad * kChunk + lane] =
          make_float4(T, cr, cg, cb);
    bool hits[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      hits[q] = false;
      const uint32_t k = round + q * kChunk + lane;
      Staged st{};
      if (k >= run.x && k < run.y) {
        stage<kCull>(splats, pair_splats, k, st, nullptr, tiles_x, tiles_y);
        hits[q] = meets(st, bx0, bx1, by0, by1) && (!kCull || holds(st.rect, wt.sx, wt.sy));
      }
      put_record(records[q], lane, st);
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const SplatRecords& rec = records[q];
      const uint32_t base = round + q * kChunk;
      uint64_t mask = __ballot(hits[q]);
      // Two splats per iteration: both alphas first (they do not depend on T), then
      // the two blends in order, exactly as one at a time.
      auto alpha_of = [&](int j, float& alpha, float& sr, float& sg, float& sb) {
        const float4 ra = rec.a[j], rc = rec.c[j];
        sr = rc.z;
        sg = rc.w;
        sb = rec.b[j].x;
        const float power2 = splat_power2(cx - ra.x, cy - ra.y, ra.z, ra.w, rc.x);
        const float a = fminf(0.99f, rc.y * __builtin_amdgcn_exp2f(power2));
        return power2 <= 0.f && a >= 1.f / 255.f ? (alpha = a, true) : (alpha = 0.f, false);
      };
      // A splat that would take T under 1e-4 is not blended and ends the pixel: its
      // weight is zeroed rather than each update selected, and T is negated.
      auto blend = [&](int j, float alpha, float sr, float sg, float sb) {
        const float remaining = T * (1.f - alpha);
        const bool stop = remaining < 0.0001f;
        const float w = stop ? 0.f : alpha * T;
        cr += sr * w;
        cg += sg * w;
        cb += sb * w;
        T = stop ? -T : remaining;
        last = stop ? last : base + j - run.x + 1;
      };
      while (mask) {
        const int j0 = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int j1 = mask ? __builtin_ctzll(mask) : j0;
        const bool second = mask != 0;
        mask &= mask - 1;
        float a0, r0, g0, b0, a1, r1, g1, b1;
        const bool v0 = alpha_of(j0, a0, r0, g0, b0);
        const bool v1 = alpha_of(j1, a1, r1, g1, b1) && second;
        if (T > 0.f && v0) blend(j0, a0, r0, g0, b0);
        if (T > 0.f && v1) blend(j1, a1, r1, g1, b1);
      }
    }
    wave_sync();  // every lane has read this round's records before the next round writes
  }
  T = fabsf(T);
  if (inside) {
    const float r = cr + T * bg_r, g = cg + T * bg_g, b = cb + T * bg_b;
    const uint32_t index = py * width + px;
    pixel_color[index] = make_float4(r, g, b, T);
    pixel_last[index] = last;
    if (target) {
      auto u8 = [](float v) { return static_cast<uint32_t>(fminf(fmaxf(v, 0.f), 1.f) * 255.f + 0.5f); };
      target[index] = u8(r) | (u8(g) << 8) | (u8(b) << 16) | (255u << 24);
    }
  }
  return ckpt ? wave_max(last) : 0u;
}

// The segmented backward's items of a tile whose list is `run` and whose last
// blended splat is `most` pairs in (see Segments): one thread's work. Only a list's
// first and last parts can be short, so a tile adds at most two of them.
__device__ inline void emit_segments(const Segments& seg, uint32_t tile, uint2 run, uint32_t most) {
  if (most == 0) return;
  const uint32_t size = 1u << seg.shift, end = run.x + most;
  const uint32_t first = run.x >> seg.shift, final = (end - 1) >> seg.shift;
  auto length = [&](uint32_t b) { return min(end, (b << seg.shift) + size) - max(run.x, b << seg.shift); };
  uint2* items = reinterpret_cast<uint2*>(seg.items);
  uint32_t whole = 0;
  for (uint32_t b = first; b <= final; b++) {
    const uint32_t len = length(b);
    if (len == size) {
      whole++;
    } else {
      const uint32_t c = 1 + (size - 1 - len) * kPartClasses / size;  // 1: longest
      const uint32_t k = atomicAdd(seg.counts + c, 1u);
      if (k < seg.part_room) items[seg.whole_room + (c - 1) * seg.part_room + k] = make_uint2(tile, b);
    }
  }
  if (whole == 0) return;
  uint32_t w = atomicAdd(seg.counts, whole);
  for (uint32_t b = first; b <= final; b++)
    if (length(b) == size) {
      if (w < seg.whole_room) items[w] = make_uint2(tile, b);
      w++;
    }
}

// One 16x16 tile (tile_x, tile_y) by a 256-thread workgroup: wave w takes quadrant w.
template <int kPer>
__device__ __forceinline__ void rasterize_tile(const float4* __restrict__ splats,
                                               const uint32_t* __restrict__ pair_splats, uint2 run, uint32_t tile_x,
                                               uint32_t tile_y, uint32_t width, uint32_t height, float bg_r,
                                               float bg_g, float bg_b, uint32_t* __restrict__ target,
                                               float4* __restrict__ pixel_color, uint32_t* __restrict__ pixel_last,
                                               const Segments& seg = Segments{}) {
  __shared__ SplatRecords s_records[4][kPer];  // one slot per wave: waves never share records
  __shared__ uint32_t s_most[4];
  const uint32_t wave = threadIdx.x / kWave;
  float4* ckpt = seg.shift ? reinterpret_cast<float4*>(seg.ckpt) : nullptr;
  const uint32_t most =
      rasterize_wave<kPer, false>(splats, pair_splats, run, WaveTile{tile_x, tile_y, wave}, width, height, bg_r,
                                  bg_g, bg_b, target, pixel_color, pixel_last, s_records[wave], ckpt, seg.shift);
  if (seg.shift) {
    if ((threadIdx.x & (kWave - 1)) == 0) s_most[wave] = most;
    __syncthreads();
    if (threadIdx.x == 0)
      emit_segments(seg, tile_y * ceil_div(width, kTile) + tile_x, run,
                    max(max(s_most[0], s_most[1]), max(s_most[2], s_most[3])));
  }
}

template <int kCtrl>
__device__ inline float dpp_add(float v) {
  return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, 0xf, 0xf,
                                                                   false));
}

// Sum over each 16-lane row with DPP; every lane of the row gets the row total.
__device__ inline float row_total(float v) {
  v = dpp_add<0xb1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4e>(v);   // quad_perm [2,3,0,1]: quad sums
  v = dpp_add<0x141>(v);  // row_half_mirror: 8-lane sums
  v = dpp_add<0x140>(v);  // row_mirror: row sums
  return v;
}

// The sum of the two results of a permlane swap. hipcc (ROCm 7.2) folds a float
// add of the swap's two results into twice the first; an empty asm barrier
// between them keeps them distinct (it emits no instruction).
__device__ inline float swap_sum(decltype(__builtin_amdgcn_permlane32_swap(0u, 0u, false, false)) r) {
  uint32_t r0 = r[0], r1 = r[1];
  __asm__ volatile("" : "+v"(r0), "+v"(r1));
  return __builtin_bit_cast(float, r0) + __builtin_bit_cast(float, r1);
}

// Four splats' per-lane values (a, b, c, d) = v[0..3][q] to four wave totals at once,
// for each of the kGrads gradients q.
// v_permlane32_swap(x, y) returns (x.lo y.lo, x.hi y.hi) by 32-lane halves and
// v_permlane16_swap(x, y) (x.r0 y.r0 x.r2 y.r2, x.r1 y.r1 x.r3 y.r3) by 16-lane
// rows: swapping halves and adding leaves a|c and b|d by half, swapping rows and
// adding leaves a|b|c|d by row, and one DPP row reduction finishes all four, so
// row r of the wave ends with splat r's total. The gradients go one step at a time
// across all of them: each cross-lane step then has eight independent ones between
// it and the step it depends on, which fill the DPP and permlane read-after-write
// wait states that one gradient's chain alone pads with s_nop.
template <int kTerms>
__device__ inline void four_totals_all(const float (&v)[4][kTerms], float (&total)[kTerms]) {
  float lo[kTerms], hi[kTerms];
#pragma unroll
  for (int q = 0; q < kTerms; q++)
    lo[q] = swap_sum(__builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, v[0][q]),
                                                      __builtin_bit_cast(uint32_t, v[2][q]), false, false));
#pragma unroll
  for (int q = 0; q < kTerms; q++)
    hi[q] = swap_sum(__builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, v[1][q]),
                                                      __builtin_bit_cast(uint32_t, v[3][q]), false, false));
#pragma unroll
  for (int q = 0; q < kTerms; q++)
    total[q] = swap_sum(__builtin_amdgcn_permlane16_swap(__builtin_bit_cast(uint32_t, lo[q]),
                                                         __builtin_bit_cast(uint32_t, hi[q]), false, false));
#pragma unroll
  for (int q = 0; q < kTerms; q++) total[q] = dpp_add<0xb1>(total[q]);
#pragma unroll
  for (int q = 0; q < kTerms; q++) total[q] = dpp_add<0x4e>(total[q]);
#pragma unroll
  for (int q = 0; q < kTerms; q++) total[q] = dpp_add<0x141>(total[q]);
#pragma unroll
  for (int q = 0; q < kTerms; q++) total[q] = dpp_add<0x140>(total[q]);
}

// The per-pixel replay state of the backward pass (rasterizeBackward.wgsl:161-177).
// The colour behind a splat enters its gradient only through its dot product with
// dL/dpixel, so the replay carries that dot product (`behind`) rather than the
// colour. It also counts the background in: with B_i the colour behind splat i
// seen from just behind it, background included, dL/dalpha_i = T_i (c_i - B_i) . g
// holds without the reference's separate -T_final (bg . g) / (1 - alpha_i) term,
// B starts at bg . g behind the last blended splat, and B_{i-1} = B_i +
// alpha_i (c_i - B_i) follows right after splat i's gradient, sharing its c_i - B_i
// (the reference defers that update by one splat and keeps the last splat's alpha
// and colour for it): about 5 fewer VALU instructions per splat and pixel.
struct Replay {
  float T, gr, gg, gb;
  float behind;  // B . g
  float loss;    // the pixel's loss share, for the footprint terms
};

// With kFoot the replay also yields two footprint terms per (splat, pixel): the
// blend weight alpha T times the pixel's loss share, and alpha T. Summed over a
// splat's pixels and views they give the loss over the Gaussian's footprint.
template <bool kFoot>
constexpr int kTermsOf = kGrads + (kFoot ? 2 : 0);

// One splat's 9 gradient terms at one pixel (rasterizeBackward.wgsl:199-230),
// written without branches: a pixel the splat does not blend into (power > 0,
// alpha < 1/255, or past the pixel's last splat) runs the same instructions with
// alpha = 0, which leaves T and the colour behind as they were and zeroes every
// term. The terms are linear in the splat's constants, so the conic and the
// opacity are applied once per splat after the reduction: with
// g = dL/dpower = o dL/dopacity (o the opacity) they are dL/dopacity times dx, dy
// (mean2: o a, b, c times these), dx², dx dy, dy² (conic: o times -1/2, -1, -1/2),
// then dL/dopacity and dL/drgb.
template <bool kFoot = false>
__device__ inline bool replay_splat(Replay& st, bool blended, float cx, float cy, float mx, float my, float ha,
                                    float hb, float hc, float op, float sr, float sg, float sb,
                                    float (&share)[kTermsOf<kFoot>]) {
  const float dx = cx - mx, dy = cy - my;
  const float power2 = splat_power2(dx, dy, ha, hb, hc);
  const float G = __builtin_amdgcn_exp2f(power2);
  const float clamped = fminf(0.99f, op * G);
  const bool valid = blended && power2 <= 0.f && clamped >= 1.f / 255.f;
  const float alpha = valid ? clamped : 0.f;
  st.T *= __builtin_amdgcn_rcpf(1.f - alpha);  // T in front of the splat; unchanged when idle
  const float color = sr * st.gr + sg * st.gg + sb * st.gb;  // this splat's colour . dL/dpixel
  const float d = color - st.behind;
  const float alpha_grad = st.T * d;
  st.behind += alpha * d;
  const float opacity_grad = valid ? alpha_grad * G : 0.f;
  const float gdx = opacity_grad * dx, gdy = opacity_grad * dy;  // x opacity after the reduction
  share[0] = gdx;
  share[1] = gdy;
  share[2] = gdx * dx;
  share[3] = gdx * dy;
  share[4] = gdy * dy;
  share[5] = opacity_grad;
  const float at = alpha * st.T;
  share[6] = at * st.gr;
  share[7] = at * st.gg;
  share[8] = at * st.gb;
  if constexpr (kFoot) {
    share[9] = at * st.loss;
    share[10] = at;
  }
  return valid;
}

// rasterizeBackward.wgsl:49-163: each pixel replays its list back to front from
// its last blended splat. A wave takes the splats that touch its quadrant four at
// a time, reduces their 9 gradients across its 64 pixels together, and its row
// leaders park the totals in LDS; per 64-splat chunk the workgroup's waves add
// theirs and one float atomic per (splat, list, gradient) goes out.
// The workgroup's LDS for one backward list with kPer x 64 pairs per round and
// kWaves waves (kCull: each staged splat's 16-px tile rectangle too).
template <int kPer, int kWaves = 4, bool kCull = false, bool kFoot = false>
struct BackwardShared {
  float sums[kWaves][kTermsOf<kFoot>][kPer * kChunk];
  float4 conic[kPer * kChunk];
  float4 box[kPer * kChunk];   // each splat's alpha >= 1/255 box (x0, x1, y0, y1)
  uint4 rect[kCull ? kPer * kChunk : 1];
  SplatRecords records[kPer];  // the round's splats, staged once for all waves
  uint32_t index[kPer * kChunk];
  uint32_t last[kWaves];
};

// The part [lo, hi) of a pair list that a backward workgroup replays (absolute pair
// positions). With ckpt, hi is a segment boundary inside the list (or the list's
// end) and a pixel whose last blended splat lies at or past hi starts its replay at
// hi from the forward pass's checkpoint there: T = the checkpoint's T, and the
// colour behind = (blended colour - checkpoint colour) / T, the colour of every
// splat from hi on as seen through the splats in front of hi. A pixel that ends
// inside the part starts from its own end as for the whole list.
struct Part {
  uint32_t lo, hi;
  const float4* __restrict__ ckpt;  // nullptr: the whole list, no checkpoint
  uint32_t shift;
};

// One pair list (or the part `part` of it) of the backward pass by a workgroup of
// kWaves waves, wave w on its quadrant `wt` (kCull: `run` is a 32-px bin's list,
// see WaveTile), kPer x 64 pairs per round; the workgroup may call it again for
// another list once it returns (it ends on a barrier). kFoot: also adds each
// splat's two footprint terms (see Replay) into footprint[2 i], footprint[2 i + 1],
// reading the per-pixel loss shares.
template <int kPer, int kWaves, bool kCull, bool kFoot = false>
__device__ __forceinline__ void rasterize_backward_group(const float4* __restrict__ splats,
                                                const uint32_t* __restrict__ pair_splats, uint2 run, WaveTile wt,
                                                uint32_t width, uint32_t height, float bg_r, float bg_g, float bg_b,
                                                const float4* __restrict__ pixel_color,
                                                const uint32_t* __restrict__ pixel_last,
                                                const float4* __restrict__ pixel_grads,
                                                float* __restrict__ splat_grads,
                                                BackwardShared<kPer, kWaves, kCull, kFoot>& sh,
                                                const float* __restrict__ pixel_loss = nullptr,
                                                float* __restrict__ footprint = nullptr,
                                                Part part = Part{0u, 0u, nullptr, 0u}) {
  constexpr int kRound = kPer * kChunk;
  constexpr int kTerms = kTermsOf<kFoot>;
  constexpr int kGroupThreads = kWaves * kWave;
  auto& s_sums = sh.sums;
  auto& s_index = sh.index;
  auto& s_last = sh.last;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t qx = wt.sx * kTile + (wt.quad & 1) * 8, qy = wt.sy * kTile + (wt.quad >> 1) * 8;
  const uint32_t px = qx + (lane & 7), py = qy + (lane >> 3);
  const uint32_t tiles_x = ceil_div(width, kTile), tiles_y = ceil_div(height, kTile);
  const bool inside = px < width && py < height;
  const float cx = static_cast<float>(px) + 0.5f, cy = static_cast<float>(py) + 0.5f;
  const float bx0 = static_cast<float>(qx) + 0.5f, bx1 = bx0 + 7.f;
  const float by0 = static_cast<float>(qy) + 0.5f, by1 = by0 + 7.f;
  Replay st{};
  uint32_t last = 0;
  float final_T = 0.f;
  float4 color = make_float4(0.f, 0.f, 0.f, 0.f);
  if (inside) {
    const uint32_t index = py * width + px;
    color = pixel_color[index];
    final_T = color.w;
    last = pixel_last[index];
    const float4 pg = pixel_grads[index];
    st.gr = pg.x; st.gg = pg.y; st.gb = pg.z;
    if (kFoot) st.loss = pixel_loss[index];
  }
  st.T = final_T;
  st.behind = bg_r * st.gr + bg_g * st.gg + bg_b * st.gb;  // behind the last blended splat: the background
  // The replayed positions [lo, hi), relative to the list.
  const uint32_t lo = part.ckpt ? part.lo - run.x : 0u;
  const uint32_t hi = part.ckpt ? part.hi - run.x : run.y - run.x;
  if (part.ckpt && last > hi) {
    const float4 c = part.ckpt[static_cast<size_t>(part.hi >> part.shift) * (4 * kChunk) + wt.quad * kChunk + lane];
    // The pixel's colour (background included) less what the splats in front of hi
    // gave, through their transmittance T: everything behind them.
    st.T = c.x;
    st.behind = ((color.x - c.y) * st.gr + (color.y - c.z) * st.gg + (color.z - c.w) * st.gb) / c.x;
    last = hi;
  }
  const uint32_t my_last = wave_max(last);
  if (lane == 0) s_last[wave] = my_last;
  __syncthreads();
  uint32_t tile_last = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) tile_last = max(tile_last, s_last[w]);
  const int row = lane >> 4;
  const bool row_leader = (lane & 15) == 0;

  const uint32_t end = run.x + hi;
  const int rounds = tile_last > lo ? static_cast<int>(ceil_div(tile_last - lo, kRound)) : 0;
  for (int c = rounds - 1; c >= 0; c--) {
    const uint32_t round = run.x + lo + static_cast<uint32_t>(c) * kRound;
    // The round's splats are staged once for every wave, 64 per staging wave.
    for (int q = wave; q < kPer; q += kWaves) {
      const uint32_t k = round + q * kChunk + lane;
      Staged st{};
      float4 conic = make_float4(0.f, 0.f, 0.f, 0.f);
      if (k < end) stage<kCull>(splats, pair_splats, k, st, &conic, tiles_x, tiles_y);
      s_index[q * kChunk + lane] = st.index;
      sh.conic[q * kChunk + lane] = conic;
      sh.box[q * kChunk + lane] = make_float4(st.x0, st.x1, st.y0, st.y1);
      if (kCull) sh.rect[q * kChunk + lane] = make_uint4(st.rect[0], st.rect[1], st.rect[2], st.rect[3]);
      put_record(sh.records[q], lane, st);
    }
#pragma unroll
    for (int q = 0; q < kPer; q++)
#pragma unroll
      for (int v = 0; v < kTerms; v++) s_sums[wave][v][q * kChunk + lane] = 0.f;
    __syncthreads();  // the staged records are read by every wave
    bool hits[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const uint32_t k = round + q * kChunk + lane;
      const float4 b = sh.box[q * kChunk + lane];
      hits[q] = k < end && k - run.x < my_last && b.y >= bx0 && b.x <= bx1 && b.w >= by0 && b.z <= by1;
      if (kCull) {
        const uint4 r = sh.rect[q * kChunk + lane];
        hits[q] = hits[q] && wt.sx >= r.x && wt.sx < r.z && wt.sy >= r.y && wt.sy < r.w;
      }
    }
    // Back to front: the last 64 of the round first.
#pragma unroll
    for (int q = kPer - 1; q >= 0; q--) {
      const SplatRecords& rec = sh.records[q];
      const uint32_t base = round + q * kChunk;
      uint64_t mask = __ballot(hits[q]);
      while (mask) {
        // Up to four splats, back to front; missing ones contribute zeros.
        float share[4][kTerms];
        int js[4];
        bool any = false;
#pragma unroll
        for (int g = 0; g < 4; g++) {
          js[g] = -1;
          bool blended = false;
          float mx = 0.f, my = 0.f, ha = 0.f, hb = 0.f, hc = 0.f, op = 0.f, sr = 0.f, sg = 0.f, sb = 0.f;
          if (mask) {
            const int j = kWave - 1 - __builtin_clzll(mask);
            mask &= ~(1ull << j);
            js[g] = q * kChunk + j;
            const float4 ra = rec.a[j], rc = rec.c[j];
            mx = ra.x; my = ra.y; ha = ra.z; hb = ra.w;
            hc = rc.x; op = rc.y; sr = rc.z; sg = rc.w; sb = rec.b[j].x;
            blended = base + j - run.x < last;
          }
          any |= replay_splat<kFoot>(st, blended, cx, cy, mx, my, ha, hb, hc, op, sr, sg, sb, share[g]);
        }
        if (!__any(any)) continue;
        const int mine = row == 0 ? js[0] : (row == 1 ? js[1] : (row == 2 ? js[2] : js[3]));
        float total[kTerms];
        four_totals_all<kTerms>(share, total);
        if (row_leader && mine >= 0) {
#pragma unroll
          for (int v = 0; v < kTerms; v++) s_sums[wave][v][mine] = total[v];
        }
      }
    }
    __syncthreads();
    // Per (slot, gradient pair): the waves' sums, then the splat's constants, each of
    // the first five sums times the opacity (g = o dL/dopacity): mean2 from the
    // (g dx, g dy) sums through the conic, the conic terms scaled by -1/2, -1, -1/2;
    // the rest as summed.
    constexpr int kParts = kFoot ? 6 : 5;
    for (int e = tid; e < kParts * kRound; e += kGroupThreads) {
      const int part = e / kRound, slot = e % kRound;
      auto sum = [&](int v) {
        float total = 0.f;
#pragma unroll
        for (int w = 0; w < kWaves; w++) total += s_sums[w][v][slot];
        return total;
      };
      const float op = sh.records[slot / kChunk].c[slot % kChunk].y;
      float* out = splat_grads + static_cast<size_t>(s_index[slot]) * kGrads;
      float t0, t1;
      int v0, v1;
      if (part == 0) {
        const float4 k = sh.conic[slot];
        const float u = sum(0) * op, w = sum(1) * op;
        t0 = k.x * u + k.y * w;
        t1 = k.y * u + k.z * w;
        v0 = 0; v1 = 1;
      } else if (part == 1) {
        t0 = -0.5f * sum(2) * op; t1 = -sum(3) * op;
        v0 = 2; v1 = 3;
      } else if (part == 2) {
        t0 = -0.5f * sum(4) * op; t1 = sum(5);
        v0 = 4; v1 = 5;
      } else if (part == 3) {
        t0 = sum(6); t1 = sum(7);
        v0 = 6; v1 = 7;
      } else if (part == 4 || !kFoot) {
        t0 = sum(8); t1 = 0.f;
        v0 = 8; v1 = -1;
      } else {  // part 5: the footprint terms
        t0 = sum(kFoot ? kGrads : 0); t1 = sum(kFoot ? kGrads + 1 : 0);
        out = footprint + static_cast<size_t>(s_index[slot]) * 2;
        v0 = 0; v1 = 1;
      }
      if (t0 != 0.f) unsafeAtomicAdd(out + v0, t0);
      if (v1 >= 0 && t1 != 0.f) unsafeAtomicAdd(out + v1, t1);
    }
    __syncthreads();
  }
  __syncthreads();  // s_last is rewritten by the next list
}

// One 16x16 tile of the backward pass by a 256-thread workgroup (wave w on
// quadrant w), kPer x 64 pairs per round.
template <int kPer>
__device__ __forceinline__ void rasterize_backward_tile(const float4* __restrict__ splats,
                                               const uint32_t* __restrict__ pair_splats, uint2 run, uint32_t tile_x,
                                               uint32_t tile_y, uint32_t width, uint32_t height, float bg_r,
                                               float bg_g, float bg_b, const float4* __restrict__ pixel_color,
                                               const uint32_t* __restrict__ pixel_last,
                                               const float4* __restrict__ pixel_grads, float* __restrict__ splat_grads,
                                               BackwardShared<kPer>& sh) {
  rasterize_backward_group<kPer, 4, false>(splats, pair_splats, run,
                                           WaveTile{tile_x, tile_y, static_cast<uint32_t>(threadIdx.x / kWave)},
                                           width, height, bg_r, bg_g, bg_b, pixel_color, pixel_last, pixel_grads,
                                           splat_grads, sh);
}

}  // namespace raster
}  // namespace dio
