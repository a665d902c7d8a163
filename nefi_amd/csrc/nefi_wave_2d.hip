// Fast 2-D float32 kernels of libnefi_wave (sponge, Cerjan and no absorbing layer; order 2/4).
//
// Mapping: a workgroup is 4 wavefronts; lane l of a wavefront owns the four columns
// x = 4*(64*blockIdx.x + l) .. +3 (one 16-byte vector, a wavefront covers 256 contiguous
// columns = 1 KiB per row: fully coalesced) and wavefront w marches down its own strip of TY
// rows. The y-part of the stencil comes from a sliding register window of 2R+1 rows; the
// x-part from the neighbouring lanes' vectors through wavefront shuffles, with lanes 0 and 63
// loading their two outer columns themselves. blockIdx.z selects a group of shots that the
// workgroup advances one after the other, so the shared rows of m (and, in the reverse sweep,
// the float64 gradient tile of the group) are re-read from cache rather than from HBM.
//
// Arithmetic: the same terms as the generic kernels (nefi_wave.hip). Implicit contraction of
// multiplies and adds is off in this file and every fused multiply-add is written out, so the
// rounding of each kernel is fixed by its source: the tile variants below (register march, LDS
// tiles, temporal blocks, damped and layer-free tiles) give bit-identical cell values.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "nefi_wave.h"
#include "nefi_wave_internal.h"

#pragma clang fp contract(off)

namespace nw {

namespace {

__device__ __forceinline__ float4 ld4(const float* p) {
  return *reinterpret_cast<const float4*>(p);
}
// Nontemporal loads for values read once per launch at a point (the oldest level of the
// leapfrog): the streaming hint keeps them from displacing the tiles that neighbouring
// workgroups read again as halos (measured: reverse steps 260-269 -> 234-235 us at 1064^2 x 64
// shots; in the blocked forward kernel, whose regions overlap, it was 1 % slower and is not used).
__device__ __forceinline__ float4 ld4nt(const float* p) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
  return make_float4(t.x, t.y, t.z, t.w);
}
// Wavefield stores are nontemporal: no kernel reads back what it writes, and the next reader
// is another launch, so the lines need not stay in the caches (measured: 5-6 % faster
// gradients at 1064^2 x 64 shots, sponge; 2 % with PML).
__device__ __forceinline__ void st4(float* p, const float4& v) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v t = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(t, reinterpret_cast<f4v*>(p));
}
__device__ __forceinline__ void st4_lds(float* p, const float4& v) {
  *reinterpret_cast<float4*>(p) = v;
}
// Plain wavefield store, for the 3-D kernels: a 3-D photoacoustic field (212^3, one shot, 38 MB)
// stays in the 256 MiB last-level cache from one launch to the next, and the nontemporal hints
// cost that workload 7 % (88.9 -> 95.7 ms per forward + adjoint). A size-dependent choice made
// at run time inside the kernels broke the 2-D kernels' load scheduling, so the choice is per
// kernel family: 2-D plans at the benchmark scale (64 shots, 290 MB per field) stream.
__device__ __forceinline__ void st4p(float* p, const float4& v) {
  *reinterpret_cast<float4*>(p) = v;
}
__device__ __forceinline__ float4 f4(float a) { return make_float4(a, a, a, a); }

// Read-only table value at a wave-uniform index (profiles, source taps): through the constant
// address space the compiler may use a scalar load, which waits only on the scalar cache, not
// on the wave's outstanding vector memory operations.
template <typename T>
__device__ __forceinline__ T ldc(const T* p, int i) {
  typedef const __attribute__((address_space(4))) T* cptr;
  return ((cptr)p)[i];
}

// Index of this lane's wavefront in its workgroup as a wave-uniform value: rows addressed
// through it are the same for every lane, so their per-row profile values (damping, PML
// coefficients) are read with scalar loads rather than vector loads that wait on memory.
__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// row y of a grid (zero outside the grid or for inactive lanes)
// offsets within one shot's grid fit in 32 bits (nw_plan_create rejects >= 2^31 cells), which
// lets loads use a uniform base address plus a 32-bit per-lane offset
__device__ __forceinline__ float4 row4(const float* g, int y, int D1, int P2, int x, bool act) {
  return (act && y >= 0 && y < D1) ? ld4(g + (y * P2 + x)) : f4(0.f);
}

// one component of the cross stencil, in the generic kernels' order (axis 1, then axis 2)
template <int R>
__device__ __forceinline__ float stencil1(const Geo<float>& g, float c, float u1, float d1,
                                          float u2, float d2, float l1, float r1, float l2,
                                          float r2) {
  float s = __builtin_fmaf(g.c1[1], u1 + d1, g.cc * c);
  if (R == 2) s = __builtin_fmaf(g.c2[1], u2 + d2, s);
  s = __builtin_fmaf(g.c1[2], l1 + r1, s);
  if (R == 2) s = __builtin_fmaf(g.c2[2], l2 + r2, s);
  return s;
}

// stencil of the centre row w[R] of a window, for the four columns of the lane
template <int R>
__device__ __forceinline__ float4 stencil4(const Geo<float>& g, const float4* w, float l2,
                                           float l1, float r1, float r2) {
  const float4& c = w[R];
  const float4 u1 = w[R - 1], d1 = w[R + 1];
  const float4 u2 = R == 2 ? w[0] : f4(0.f), d2 = R == 2 ? w[2 * R] : f4(0.f);
  float4 out;
  out.x = stencil1<R>(g, c.x, u1.x, d1.x, u2.x, d2.x, l1, c.y, l2, c.z);
  out.y = stencil1<R>(g, c.y, u1.y, d1.y, u2.y, d2.y, c.x, c.z, l1, c.w);
  out.z = stencil1<R>(g, c.z, u1.z, d1.z, u2.z, d2.z, c.y, c.w, c.x, r1);
  out.w = stencil1<R>(g, c.w, u1.w, d1.w, u2.w, d2.w, c.z, r1, c.y, r2);
  return out;
}

// damping of one cell from the row term hy and the column term hx
template <int ABS>
__device__ __forceinline__ void damp1(float hy, float hx, float& a, float& b, float& tap) {
  a = 1.f;
  b = 1.f;
  tap = 1.f;
  if (ABS == NW_SPONGE || ABS == NW_PML) {
    const float half = hy + hx;
    a = 1.f / (1.f + half);
    b = 1.f - half;
  } else if (ABS == NW_CERJAN) {
    tap = hy < 1.f ? hy : 1.f;
    tap = hx < tap ? hx : tap;
    a = tap;
  }
}

}  // namespace

// ------------------------------------------------------------------------------------------
// tile mapping
// ------------------------------------------------------------------------------------------
// MI355X places workgroup i of a dispatch on XCD i % 8 and every XCD has its own L2. Tiles are
// numbered column chunk fastest, then row strip, then shot group, and XCD j is given the
// contiguous range [j*per, (j+1)*per) of that order, so the strips above and below a tile
// (whose rows form its halo) run on the same XCD and share its L2.
constexpr int kXcds = 8;

struct Tile {
  int cx, ry, grp;
  bool valid;
};

__device__ __forceinline__ Tile tile_of(int nX, int nY, int nG, int xcd_map) {
  const int T = nX * nY * nG;
  const int per = (T + kXcds - 1) / kXcds;
  const int i = blockIdx.x;
  const int t = xcd_map ? (i % kXcds) * per + i / kXcds : i;
  Tile r;
  r.valid = t < T;
  r.cx = t % nX;
  r.ry = (t / nX) % nY;
  r.grp = t / (nX * nY);
  return r;
}

__host__ __device__ inline int div_up(int a, int b) { return (a + b - 1) / b; }

// Tile columns of k4_steps: `full` columns of out_w output columns each and, when the grid's
// last columns fit in half a wavefront row (128 - 2 hx of them), one split column. A split
// workgroup steps two half-width tiles of that column, row tiles 2j (lanes 0-31) and 2j + 1
// (lanes 32-63), so the narrow rest of the grid takes half as many workgroups.
struct BlockCols {
  int full;
  bool split;
};
template <class S>
__host__ __device__ inline BlockCols block_cols(int D2) {
  const int full = D2 / S::out_w, rest = D2 - full * S::out_w;
  if (rest == 0) return {full, false};
  if (rest <= 128 - 2 * S::hx) return {full, true};
  return {full + 1, false};
}
__host__ __device__ inline int block_tiles(BlockCols c, int nY, int nG) {
  return (c.split ? (nY + 1) / 2 * (2 * c.full + 1) : c.full * nY) * nG;
}
// Tile of this workgroup for block_cols' layout, per shot group; the XCD map as in tile_of.
// With a split column, each pair of row tiles lists its full tiles and then its split tile,
// so the split workgroups are spread over the launch rather than left to its end (for an odd
// number of row tiles, the last pair's second row of full tiles is listed and returns).
__device__ __forceinline__ Tile tile_of_cols(BlockCols c, int nY, int nG, int xcd_map) {
  const int per_grp = block_tiles(c, nY, 1);
  const int T = per_grp * nG;
  const int per = (T + kXcds - 1) / kXcds;
  const int i = blockIdx.x;
  const int t = xcd_map ? (i % kXcds) * per + i / kXcds : i;
  Tile r;
  r.grp = t / per_grp;
  const int u = t - r.grp * per_grp;
  if (!c.split) {
    r.cx = u % c.full;
    r.ry = u / c.full;
  } else {
    const int pair = u / (2 * c.full + 1), v = u - pair * (2 * c.full + 1);
    if (v < 2 * c.full) {
      r.cx = v % c.full;
      r.ry = 2 * pair + v / c.full;
    } else {
      r.cx = c.full;
      r.ry = 2 * pair;
    }
  }
  r.valid = t < T && r.ry < nY;
  return r;
}
// Tile columns of k3_adjoint (256 output columns each): the last 128 or fewer columns of the
// grid are stepped as split workgroups, two half-width tiles each.
__host__ __device__ inline BlockCols adj_cols(int D2) {
  const int full = D2 / 256, rest = D2 - full * 256;
  if (rest == 0) return {full, false};
  if (rest <= 128) return {full, true};
  return {full + 1, false};
}

__device__ __forceinline__ float2 ld2(const float* p) {
  return *reinterpret_cast<const float2*>(p);
}

// the two outer columns lanes 0 and 63 need (x-2, x-1 for lane 0; x+4, x+5 for lane 63),
// as one 8-byte load per row; zero outside the grid and for the other lanes
__device__ __forceinline__ float2 edge2(const float* g, int y, int D2, int P2, int x, int lane) {
  const float* rp = g + y * P2;
  if (lane == 0 && x >= 4) return ld2(rp + x - 2);
  if (lane == 63 && x + 4 < D2) return ld2(rp + x + 4);
  return make_float2(0.f, 0.f);
}

// x-neighbours of the lane's four columns: (x-2, x-1, x+4, x+5) from the neighbouring lanes,
// with lanes 0 and 63 taking them from their edge pair. Returned by value and chosen with
// selects so the compiler keeps them in registers.
__device__ __forceinline__ float4 xnb(const float4& v, const float2& e, int lane) {
  const float l2 = __shfl_up(v.z, 1, 64);
  const float l1 = __shfl_up(v.w, 1, 64);
  const float r1 = __shfl_down(v.x, 1, 64);
  const float r2 = __shfl_down(v.y, 1, 64);
  return make_float4(lane == 0 ? e.x : l2, lane == 0 ? e.y : l1, lane == 63 ? e.x : r1,
                     lane == 63 ? e.y : r2);
}

// ------------------------------------------------------------------------------------------
// per-cell arithmetic on the four columns of a lane (plain float4 values: nothing indexable,
// so nothing is placed in scratch memory)
// ------------------------------------------------------------------------------------------
template <int ABS>
struct Damp4 {
  float4 a, b, tap;
};

template <int ABS>
__device__ __forceinline__ Damp4<ABS> damp4(float hy, const float4& hx) {
  Damp4<ABS> d;
  damp1<ABS>(hy, hx.x, d.a.x, d.b.x, d.tap.x);
  damp1<ABS>(hy, hx.y, d.a.y, d.b.y, d.tap.y);
  damp1<ABS>(hy, hx.z, d.a.z, d.b.z, d.tap.z);
  damp1<ABS>(hy, hx.w, d.a.w, d.b.w, d.tap.w);
  return d;
}

// rhs = (2 p - b p_prev~) + m lap, with p_prev~ = tap * p_prev for a tapered previous state
template <int ABS>
__device__ __forceinline__ float rhs1(float pc, float prev_raw, float m, float lap, float b,
                                      float tap, int taper_prev) {
  const float prev = (ABS == NW_CERJAN && taper_prev) ? tap * prev_raw : prev_raw;
  return __builtin_fmaf(m, lap, __builtin_fmaf(-b, prev, 2.f * pc));
}

// adjoint of one cell: lambda^n from rbar = a lambda^{n+1}, the next step's lambda^{n+2} term
// and the stencil of q
template <int ABS>
__device__ __forceinline__ float adj1(float lc, float l2c, float lapq, float a, float b,
                                      float tap, int taper_next, float& rbar) {
  rbar = a * lc;
  const float tp1 = (ABS == NW_CERJAN && taper_next) ? tap : 1.f;
  const float pbar_next = -(b * tp1) * (a * l2c);
  return __builtin_fmaf(2.f, rbar, pbar_next) + lapq;
}

__device__ __forceinline__ float4 mul4(const float4& a, const float4& b) {
  return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
}

// Damping of a lane's four columns in a row whose row term is zero, computed once per lane:
// with hy = 0 the sponge's half = hy + hx is hx exactly, so a = 1 / (1 + hx) and b = 1 - hx
// are the per-cell values bit for bit, and the division leaves the per-row work.
template <int ABS>
__device__ __forceinline__ Damp4<ABS> damp4_col(const float4& hx) {
  return damp4<ABS>(0.f, hx);
}

// damping of one row: the precomputed column values when the row term vanishes (sponge)
template <int ABS>
__device__ __forceinline__ Damp4<ABS> damp4_row(float hy, const float4& hx,
                                                const Damp4<ABS>& col) {
  if ((ABS == NW_SPONGE || ABS == NW_PML) && hy == 0.f) return col;
  return damp4<ABS>(hy, hx);
}

// ------------------------------------------------------------------------------------------
// forward step
// ------------------------------------------------------------------------------------------
template <int R, int ABS, int TY>
__device__ __forceinline__ void step_row(const Geo<float>& g, const StepIO<float>& io, int s,
                                         int64_t base, int y, int x, int P2, int lane, bool act,
                                         const float4& hx, float w_n, float4* w, const float4& pp,
                                         const float4& mm, const float2& ed) {
  const float4 nb = xnb(w[R], ed, lane);
  if (!act) return;
  const int ro = y * P2 + x;
  const float4 lap = stencil4<R>(g, w, nb.x, nb.y, nb.z, nb.w);
  const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
  const Damp4<ABS> dm = damp4<ABS>(hy, hx);
  const float4 pc = w[R];
  const int tpv = io.taper_prev;
  float4 o = make_float4(rhs1<ABS>(pc.x, pp.x, mm.x, lap.x, dm.b.x, dm.tap.x, tpv),
                         rhs1<ABS>(pc.y, pp.y, mm.y, lap.y, dm.b.y, dm.tap.y, tpv),
                         rhs1<ABS>(pc.z, pp.z, mm.z, lap.z, dm.b.z, dm.tap.z, tpv),
                         rhs1<ABS>(pc.w, pp.w, mm.w, lap.w, dm.b.w, dm.tap.w, tpv));
  for (int t = 0; t < io.src_taps; ++t) {
    const int d = ldc(io.src_idx, s * io.src_taps + t) - ro;
    if (d >= 0 && d < 4) {
      const float v = ldc(io.src_val, s * io.src_taps + t) * w_n;
      o.x += d == 0 ? v : 0.f;
      o.y += d == 1 ? v : 0.f;
      o.z += d == 2 ? v : 0.f;
      o.w += d == 3 ? v : 0.f;
    }
  }
  if (ABS != NW_NONE) o = mul4(dm.a, o);
  st4(io.p_next + base + ro, o);
  if (io.lap_out) st4(io.lap_out + base + ro, lap);
  if (io.ck_prev) {
    st4(io.ck_prev + base + ro, pp);
    st4(io.ck_cur + base + ro, pc);
  }
}

template <int R, int ABS, int TY, int U>
__global__ void __launch_bounds__(256) k2_step(Geo<float> g, StepIO<float> io, int shots,
                                               int spw, int xcd_map) {
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const Tile tl = tile_of(div_up(D2, 256), div_up(D1, 4 * TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int x = (tl.cx * 64 + lane) * 4;
  const int ys = (tl.ry * 4 + wave) * TY;
  if (ys >= D1) return;
  const int ye = ys + TY < D1 ? ys + TY : D1;
  const bool act = x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const float w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.f;
  const int s_end = min(shots, (tl.grp + 1) * spw);
  for (int s = tl.grp * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const float* P = io.p + base;
    const float* PP = io.p_prev + base;
    float4 w[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) w[k] = row4(P, ys - R + k, D1, P2, x, act);
    if (U == 1) {
      for (int y = ys; y < ye; ++y) {
        w[2 * R] = row4(P, y + R, D1, P2, x, act);
        const float4 pp = act ? ld4(PP + (y * P2 + x)) : f4(0.f);
        const float4 mm = act ? ld4(io.m + (y * P2 + x)) : f4(0.f);
        const float2 ed = edge2(P, y, D2, P2, x, lane);
        step_row<R, ABS, TY>(g, io, s, base, y, x, P2, lane, act, hx, w_n, w, pp, mm, ed);
#pragma unroll
        for (int k = 0; k < 2 * R; ++k) w[k] = w[k + 1];
      }
    } else {
      for (int y = ys; y < ye; y += 2) {
        const bool two = y + 1 < ye;  // uniform across the wavefront
        // every load of both rows before any use
        const float4 n0 = row4(P, y + R, D1, P2, x, act);
        const float4 n1 = two ? row4(P, y + R + 1, D1, P2, x, act) : f4(0.f);
        const float4 pp0 = act ? ld4(PP + (y * P2 + x)) : f4(0.f);
        const float4 pp1 = (act && two) ? ld4(PP + ((y + 1) * P2 + x)) : f4(0.f);
        const float4 m0 = act ? ld4(io.m + (y * P2 + x)) : f4(0.f);
        const float4 m1 = (act && two) ? ld4(io.m + ((y + 1) * P2 + x)) : f4(0.f);
        const float2 e0 = edge2(P, y, D2, P2, x, lane);
        const float2 e1 = two ? edge2(P, y + 1, D2, P2, x, lane) : make_float2(0.f, 0.f);
        w[2 * R] = n0;
        step_row<R, ABS, TY>(g, io, s, base, y, x, P2, lane, act, hx, w_n, w, pp0, m0, e0);
#pragma unroll
        for (int k = 0; k < 2 * R; ++k) w[k] = w[k + 1];
        if (two) {
          w[2 * R] = n1;
          step_row<R, ABS, TY>(g, io, s, base, y + 1, x, P2, lane, act, hx, w_n, w, pp1, m1,
                               e1);
#pragma unroll
          for (int k = 0; k < 2 * R; ++k) w[k] = w[k + 1];
        }
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// reverse step with the imaging sum
// ------------------------------------------------------------------------------------------
template <int R, int ABS, int TY>
__global__ void __launch_bounds__(256) k2_adjoint(Geo<float> g, AdjIO<float> io, int shots,
                                                  int spw, int xcd_map) {
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const Tile tl = tile_of(div_up(D2, 256), div_up(D1, 4 * TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int x = (tl.cx * 64 + lane) * 4;
  const int ys = (tl.ry * 4 + wave) * TY;
  if (ys >= D1) return;
  const int ye = ys + TY < D1 ? ys + TY : D1;
  const bool act = x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const int s_end = min(shots, (tl.grp + 1) * spw);
  // gradient slice of this shot group (one float64 accumulator grid per group)
  double* G = io.grad ? io.grad + (int64_t)tl.grp * g.cells : nullptr;
  const float* MA = io.ma;
  const int tpn = io.taper_next;
  for (int s = tl.grp * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const float* L1 = io.lam1 + base;
    const float* Pn = io.p_n ? io.p_n + base : nullptr;
    const float* Ln = io.lap_n ? io.lap_n + base : nullptr;
    // window of q = (m a_eff) * lambda^{n+1}, formed as the rows arrive
    float4 qw[2 * R + 1], pw[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) {
      qw[k] = mul4(row4(MA, ys - R + k, D1, P2, x, act), row4(L1, ys - R + k, D1, P2, x, act));
      pw[k] = Pn ? row4(Pn, ys - R + k, D1, P2, x, act) : f4(0.f);
    }
    for (int y = ys; y < ye; ++y) {
      const int ro = y * P2 + x;
      // every load of the row first
      const float4 ln = row4(L1, y + R, D1, P2, x, act);
      const float4 mn = row4(MA, y + R, D1, P2, x, act);
      const float4 pn = Pn ? row4(Pn, y + R, D1, P2, x, act) : f4(0.f);
      const float4 l2c = act ? ld4nt(io.lam2 + base + ro) : f4(0.f);
      const float4 lc = act ? ld4(L1 + ro) : f4(0.f);  // read R rows ago: a cache hit
      const float2 el = edge2(L1, y, D2, P2, x, lane);
      const float2 em = edge2(MA, y, D2, P2, x, lane);
      const float2 ep = Pn ? edge2(Pn, y, D2, P2, x, lane) : make_float2(0.f, 0.f);
      double2* gp = (G && (Pn || Ln) && act) ? reinterpret_cast<double2*>(G + ro) : nullptr;
      double2 g0 = make_double2(0.0, 0.0), g1 = g0;
      if (gp) {
        g0 = gp[0];
        g1 = gp[1];
      }
      qw[2 * R] = mul4(mn, ln);
      pw[2 * R] = pn;
      const float4 qn = xnb(qw[R], make_float2(em.x * el.x, em.y * el.y), lane);
      const float4 pnb = Pn ? xnb(pw[R], ep, lane) : f4(0.f);
      if (act) {
        const float4 lapq = stencil4<R>(g, qw, qn.x, qn.y, qn.z, qn.w);
        const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
        const Damp4<ABS> dm = damp4<ABS>(hy, hx);
        float4 rb;
        const float4 o = make_float4(
            adj1<ABS>(lc.x, l2c.x, lapq.x, dm.a.x, dm.b.x, dm.tap.x, tpn, rb.x),
            adj1<ABS>(lc.y, l2c.y, lapq.y, dm.a.y, dm.b.y, dm.tap.y, tpn, rb.y),
            adj1<ABS>(lc.z, l2c.z, lapq.z, dm.a.z, dm.b.z, dm.tap.z, tpn, rb.z),
            adj1<ABS>(lc.w, l2c.w, lapq.w, dm.a.w, dm.b.w, dm.tap.w, tpn, rb.w));
        st4(io.lam0 + base + ro, o);
        if (gp) {
          const float4 lapp = Ln ? ld4(Ln + ro) : stencil4<R>(g, pw, pnb.x, pnb.y, pnb.z, pnb.w);
          g0.x = __builtin_fma((double)rb.x, (double)lapp.x, g0.x);
          g0.y = __builtin_fma((double)rb.y, (double)lapp.y, g0.y);
          g1.x = __builtin_fma((double)rb.z, (double)lapp.z, g1.x);
          g1.y = __builtin_fma((double)rb.w, (double)lapp.w, g1.y);
          gp[0] = g0;
          gp[1] = g1;
        }
      }
#pragma unroll
      for (int k = 0; k < 2 * R; ++k) {
        qw[k] = qw[k + 1];
        pw[k] = pw[k + 1];
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// LDS-tiled variants: one workgroup = one tile of TY rows x 256 columns. Every tile load (with
// its R-row halo and 4-column margins) is issued at once, so a tile costs one memory round
// trip and its halo is read once; wavefront w then marches rows w*TY/4 .. (w+1)*TY/4-1 of the
// tile out of LDS. Rows of one tile are fully unrolled, so per-row values stay in registers.
// ------------------------------------------------------------------------------------------
constexpr int kLW = 256 + 8;  // LDS row: 4 margin columns on each side, 16-byte aligned

// LDS row of a split tile (two half-width tiles side by side, each with its 4 margin columns)
constexpr int kLW2 = 2 * (128 + 8);

// Workgroup barrier that orders LDS only: waits for this wavefront's LDS accesses, not for its
// outstanding global loads and stores (which __syncthreads' release fence would drain). The
// memory clobber keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <int TY, int R>
struct TileShape {
  static constexpr int rows = TY + 2 * R;
  static constexpr int n4 = rows * (kLW / 4);  // float4 elements of one tile
  static constexpr int per_thread = (n4 + 255) / 256;
  static constexpr int my_rows = TY / 4;
  static constexpr int n4s = rows * (kLW2 / 4);  // ... of a split tile
  static constexpr int per_thread_s = (n4s + 255) / 256;
};

// issue every float4 load of a tile (rows y0-R .. y0+TY+R-1, columns x0-4 .. x0+259)
template <int TY, int R>
__device__ __forceinline__ void tile_load(const float* g, int y0, int x0, int D1, int D2, int P2,
                                          float4* v) {
  using S = TileShape<TY, R>;
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int row = i / (kLW / 4), c4 = i % (kLW / 4);
    const int gy = y0 - R + row, gx = x0 - 4 + 4 * c4;
    const bool ok = i < S::n4 && gy >= 0 && gy < D1 && gx >= 0 && gx < D2;
    v[k] = ok ? ld4(g + (gy * P2 + gx)) : f4(0.f);
  }
}

template <int TY, int R>
__device__ __forceinline__ void tile_store(float* lds, const float4* v) {
  using S = TileShape<TY, R>;
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    if (i < S::n4) *reinterpret_cast<float4*>(lds + 4 * i) = v[k];
  }
}

// row r of the LDS tile at the lane's four columns, and its x-neighbours (x-2, x-1, x+4, x+5)
__device__ __forceinline__ float4 lrow(const float* lds, int r, int lane) {
  return *reinterpret_cast<const float4*>(lds + r * kLW + 4 * lane + 4);
}
__device__ __forceinline__ float4 lnb(const float* lds, int r, int lane) {
  const float2 a = *reinterpret_cast<const float2*>(lds + r * kLW + 4 * lane + 2);
  const float2 b = *reinterpret_cast<const float2*>(lds + r * kLW + 4 * lane + 8);
  return make_float4(a.x, a.y, b.x, b.y);
}

__device__ __forceinline__ Tile tile_of_rows(int nX, int nY, int nG, int xcd_map) {
  return tile_of(nX, nY, nG, xcd_map);
}

template <int R, int ABS, int TY>
__global__ void __launch_bounds__(256) k3_step(Geo<float> g, StepIO<float> io, int shots,
                                               int spw, int xcd_map) {
  using S = TileShape<TY, R>;
  __shared__ __align__(16) float lp[S::rows * kLW];
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const Tile tl = tile_of(div_up(D2, 256), div_up(D1, TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;  // uniform over the workgroup
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  const int x = x0 + 4 * lane;
  const int r0 = wave * S::my_rows;  // first tile row of this wavefront
  const bool act = x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const Damp4<ABS> dcol = damp4_col<ABS>(hx);
  const float w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.f;
  // model rows and damping terms of this wavefront: shot-independent, loaded once for the group
  float4 mm[S::my_rows];
  float hyv[S::my_rows];
#pragma unroll
  for (int k = 0; k < S::my_rows; ++k) {
    const int y = y0 + r0 + k;
    mm[k] = (act && y < D1) ? ld4(io.m + (y * P2 + x)) : f4(0.f);
    hyv[k] = (ABS != NW_NONE && y < D1) ? g.damp[1][y] : 0.f;
  }
  const int s_end = min(shots, (tl.grp + 1) * spw);
  for (int s = tl.grp * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const float* P = io.p + base;
    float4 v[S::per_thread];
    tile_load<TY, R>(P, y0, x0, D1, D2, P2, v);
    float4 pp[S::my_rows];
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      pp[k] = (act && y < D1) ? ld4nt(io.p_prev + base + (y * P2 + x)) : f4(0.f);
    }
    __syncthreads();  // the previous shot's reads of lp are done
    tile_store<TY, R>(lp, v);
    __syncthreads();
    float4 w[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) w[k] = lrow(lp, r0 + k, lane);
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      w[2 * R] = lrow(lp, r0 + k + 2 * R, lane);
      if (act && y < D1) {
        const float4 nb = lnb(lp, r0 + k + R, lane);
        const int ro = y * P2 + x;
        const float4 lap = stencil4<R>(g, w, nb.x, nb.y, nb.z, nb.w);
        const Damp4<ABS> dm = damp4_row<ABS>(hyv[k], hx, dcol);
        const float4 pc = w[R];
        const int tpv = io.taper_prev;
        float4 o = make_float4(rhs1<ABS>(pc.x, pp[k].x, mm[k].x, lap.x, dm.b.x, dm.tap.x, tpv),
                               rhs1<ABS>(pc.y, pp[k].y, mm[k].y, lap.y, dm.b.y, dm.tap.y, tpv),
                               rhs1<ABS>(pc.z, pp[k].z, mm[k].z, lap.z, dm.b.z, dm.tap.z, tpv),
                               rhs1<ABS>(pc.w, pp[k].w, mm[k].w, lap.w, dm.b.w, dm.tap.w, tpv));
        for (int t = 0; t < io.src_taps; ++t) {
          const int d = ldc(io.src_idx, s * io.src_taps + t) - ro;
          if (d >= 0 && d < 4) {
            const float sv = ldc(io.src_val, s * io.src_taps + t) * w_n;
            o.x += d == 0 ? sv : 0.f;
            o.y += d == 1 ? sv : 0.f;
            o.z += d == 2 ? sv : 0.f;
            o.w += d == 3 ? sv : 0.f;
          }
        }
        if (ABS != NW_NONE) o = mul4(dm.a, o);
        st4(io.p_next + base + ro, o);
        if (io.lap_out) st4(io.lap_out + base + ro, lap);
        if (io.ck_prev) {
          st4(io.ck_prev + base + ro, pp[k]);
          st4(io.ck_cur + base + ro, pc);
        }
      }
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) w[j] = w[j + 1];
    }
  }
}

// Waves of one k3_adjoint workgroup: 8, one tile row each at 8-row tiles. At 8 rows a wave's
// registers stay below 128, so 4 waves share each SIMD, and every load of a shot (the two tiles
// with their halos and the wave's own rows) is in flight at once.
template <int TY>
struct AdjWaves {
  static constexpr int waves = TY < 8 ? TY : 8;
};

template <int TY, int R>
struct AdjShape {
  static constexpr int waves = AdjWaves<TY>::waves;
  static constexpr int threads = 64 * waves;
  static constexpr int rows = TY + 2 * R;
  static constexpr int my_rows = TY / waves;
  static constexpr int n4 = rows * (kLW / 4);  // float4 elements of one tile
  static constexpr int per_thread = (n4 + threads - 1) / threads;
  static constexpr int n4s = rows * (kLW2 / 4);  // ... of a split tile
  static constexpr int per_thread_s = (n4s + threads - 1) / threads;
};

// One tile of k3_adjoint. Q tiles (tile, halo and margins inside the grid, output cells in the
// box of neutral damping) skip every bounds check and step with a = b = 1, as k4_tile does.
// The tile of m a_eff is shot-independent: it stays in registers for the whole shot group, and
// so do the rows' damping terms. Per shot, every load (the lambda^{n+1} and p^n tiles with
// their halos, the wave's lambda^{n+1}, lambda^{n+2} rows) is issued before the first barrier,
// so a shot costs one memory round trip; the barriers order LDS only, so the previous shot's
// stores need not have completed.
// SP ("split", never quiet): two half-width tiles of the grid's last 128 or fewer columns, row
// tiles 2j (lanes 0-31) and 2j + 1 (lanes 32-63). An LDS row (kLW2 floats) holds both tiles'
// rows, each between its own margin columns, so every lane's x-neighbours are its own tile's.
template <int R, int ABS0, int TY, bool Q, bool SP = false>
__device__ __forceinline__ void k3_adj_tile(const Geo<float>& g, const AdjIO<float>& io,
                                            int shots, int spw, const Tile& tl, float* lq,
                                            float* lp) {
  static_assert(!(Q && SP), "split tiles lie on the grid's edge");
  constexpr int ABS = Q ? NW_NONE : ABS0;
  using S = AdjShape<TY, R>;
  constexpr int NT = S::threads;
  constexpr int LW = SP ? kLW2 : kLW;                             // LDS row pitch
  constexpr int PT = SP ? S::per_thread_s : S::per_thread;        // tile loads per thread
  constexpr int N4 = SP ? S::n4s : S::n4;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int hl = SP ? (lane & 31) : lane;  // lane within its tile
  const int x0 = tl.cx * 256, yA = tl.ry * TY;
  const int y0 = SP && lane >= 32 ? yA + TY : yA;
  const int x = x0 + 4 * hl;
  // this lane's float offset in an LDS row: 4 margin floats, then its 4 cells (the lower tile
  // of a split row starts after the upper tile's right margin and its own left margin)
  const int lo = 4 * lane + 4 + (SP && lane >= 32 ? 8 : 0);
  auto row4 = [&](const float* lds, int r) {
    return *reinterpret_cast<const float4*>(lds + r * LW + lo);
  };
  auto nb4 = [&](const float* lds, int r) {
    const float2 a = *reinterpret_cast<const float2*>(lds + r * LW + lo - 2);
    const float2 b = *reinterpret_cast<const float2*>(lds + r * LW + lo + 4);
    return make_float4(a.x, a.y, b.x, b.y);
  };
  const int r0 = wave * S::my_rows;
  const bool act = Q || x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const Damp4<ABS> dcol = damp4_col<ABS>(hx);
  const bool kept_lap = io.lap_n != nullptr;
  const bool imaging = (io.p_n != nullptr || kept_lap) && io.grad != nullptr;
  const bool p_tile = imaging && !kept_lap;
  const int tpn = io.taper_next;
  // tile loads: every element inside the grid for Q tiles
  auto load_tile = [&](const float* src, float4* v) {
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int i = threadIdx.x + NT * k;
      const int row = i / (LW / 4);
      int c4 = i % (LW / 4), ty0 = yA;
      if (SP && c4 >= kLW2 / 8) {  // the lower tile's half of the LDS row
        c4 -= kLW2 / 8;
        ty0 = yA + TY;
      }
      const int gy = ty0 - R + row, gx = x0 - 4 + 4 * c4;
      const bool in = NT * (k + 1) <= N4 || i < N4;
      const bool ok = in && (Q || (gy >= 0 && gy < D1 && gx >= 0 && gx < D2));
      v[k] = ok ? ld4(src + (gy * P2 + gx)) : f4(0.f);
    }
  };
  auto store_tile = [&](float* lds, const float4* v) {
#pragma unroll
    for (int k = 0; k < PT; ++k) {
      const int i = threadIdx.x + NT * k;
      if (i < N4) *reinterpret_cast<float4*>(lds + 4 * i) = v[k];
    }
  };
  float4 vm[PT];
  load_tile(io.ma, vm);
  // damping terms of the wave's rows (a scalar per row, loaded once, not per shot)
  float hyv[S::my_rows];
#pragma unroll
  for (int k = 0; k < S::my_rows; ++k) {
    const int y = y0 + r0 + k;
    hyv[k] = (ABS != NW_NONE && act && y < D1) ? g.damp[1][y] : 0.f;
  }
  // float64 imaging sums of this wavefront's cells over the shot group
  double2 ga[S::my_rows], gb[S::my_rows];
#pragma unroll
  for (int k = 0; k < S::my_rows; ++k) {
    ga[k] = make_double2(0.0, 0.0);
    gb[k] = ga[k];
  }
  const int s_end = min(shots, (tl.grp + 1) * spw);
  for (int s = tl.grp * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const float* L1 = io.lam1 + base;
    float4 vl[PT], vp[PT];
    load_tile(L1, vl);
    if (p_tile) load_tile(io.p_n + base, vp);
    float4 lc[S::my_rows], l2c[S::my_rows], lpn[S::my_rows];
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      const bool ok = Q || (act && y < D1);
      lc[k] = ok ? ld4(L1 + (y * P2 + x)) : f4(0.f);  // also in the tile: a cache hit
      l2c[k] = ok ? ld4nt(io.lam2 + base + (y * P2 + x)) : f4(0.f);
      lpn[k] = (ok && imaging && kept_lap) ? ld4(io.lap_n + base + (y * P2 + x)) : f4(0.f);
    }
    lds_barrier();  // the previous shot's reads of lq / lp are done
#pragma unroll
    for (int k = 0; k < PT; ++k) vl[k] = mul4(vm[k], vl[k]);
    store_tile(lq, vl);
    if (p_tile) store_tile(lp, vp);
    lds_barrier();
    float4 qw[2 * R + 1], pw[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) {
      qw[k] = row4(lq, r0 + k);
      pw[k] = p_tile ? row4(lp, r0 + k) : f4(0.f);
    }
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      qw[2 * R] = row4(lq, r0 + k + 2 * R);
      pw[2 * R] = p_tile ? row4(lp, r0 + k + 2 * R) : f4(0.f);
      if (Q || (act && y < D1)) {
        const float4 qn = nb4(lq, r0 + k + R);
        const float4 lapq = stencil4<R>(g, qw, qn.x, qn.y, qn.z, qn.w);
        const Damp4<ABS> dm = damp4_row<ABS>(hyv[k], hx, dcol);
        float4 rb;
        const float4 o = make_float4(
            adj1<ABS>(lc[k].x, l2c[k].x, lapq.x, dm.a.x, dm.b.x, dm.tap.x, tpn, rb.x),
            adj1<ABS>(lc[k].y, l2c[k].y, lapq.y, dm.a.y, dm.b.y, dm.tap.y, tpn, rb.y),
            adj1<ABS>(lc[k].z, l2c[k].z, lapq.z, dm.a.z, dm.b.z, dm.tap.z, tpn, rb.z),
            adj1<ABS>(lc[k].w, l2c[k].w, lapq.w, dm.a.w, dm.b.w, dm.tap.w, tpn, rb.w));
        st4(io.lam0 + base + (y * P2 + x), o);
        if (imaging) {
          float4 lapp = lpn[k];
          if (p_tile) {
            const float4 pn = nb4(lp, r0 + k + R);
            lapp = stencil4<R>(g, pw, pn.x, pn.y, pn.z, pn.w);
          }
          ga[k].x = __builtin_fma((double)rb.x, (double)lapp.x, ga[k].x);
          ga[k].y = __builtin_fma((double)rb.y, (double)lapp.y, ga[k].y);
          gb[k].x = __builtin_fma((double)rb.z, (double)lapp.z, gb[k].x);
          gb[k].y = __builtin_fma((double)rb.w, (double)lapp.w, gb[k].y);
        }
      }
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) {
        qw[j] = qw[j + 1];
        pw[j] = pw[j + 1];
      }
    }
  }
  if (imaging) {
    double* G = io.grad + (int64_t)tl.grp * g.cells;
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      if (Q || (act && y < D1)) {
        double2* gp = reinterpret_cast<double2*>(G + (y * P2 + x));
        double2 a = gp[0], b = gp[1];
        a.x += ga[k].x;
        a.y += ga[k].y;
        b.x += gb[k].x;
        b.y += gb[k].y;
        gp[0] = a;
        gp[1] = b;
      }
    }
  }
}

template <int R, int ABS, int TY>
__global__ void __launch_bounds__(64 * AdjWaves<TY>::waves) k3_adjoint(Geo<float> g,
                                                                      AdjIO<float> io,
                                                                      int shots, int spw,
                                                                      int xcd_map) {
  using S = AdjShape<TY, R>;
  __shared__ __align__(16) float lq[S::rows * kLW2];  // q = (m a_eff) lambda^{n+1}
  __shared__ __align__(16) float lp[S::rows * kLW2];  // p^n
  const int D1 = (int)g.D[1], D2 = (int)g.D[2];
  const BlockCols cols = adj_cols(D2);
  const Tile tl = tile_of_cols(cols, div_up(D1, TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;  // uniform over the workgroup
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  const bool quiet = y0 - R >= 0 && y0 + TY + R <= D1 && x0 >= 4 && x0 + 260 <= D2 &&
                     y0 >= g.quiet[1][0] && y0 + TY <= g.quiet[1][1] &&
                     x0 >= g.quiet[2][0] && x0 + 256 <= g.quiet[2][1];
  if (cols.split && tl.cx == cols.full)
    k3_adj_tile<R, ABS, TY, false, true>(g, io, shots, spw, tl, lq, lp);
  else if (quiet)
    k3_adj_tile<R, ABS, TY, true>(g, io, shots, spw, tl, lq, lp);
  else
    k3_adj_tile<R, ABS, TY, false>(g, io, shots, spw, tl, lq, lp);
}

// ------------------------------------------------------------------------------------------
// temporally blocked forward steps
// ------------------------------------------------------------------------------------------
// One workgroup advances one tile of TY output rows x (256 - 2 HX) output columns by up to T
// steps, reading p^{n-1}, p^n and m once and writing only the levels the caller keeps. The
// loaded region spans rows y0 - HY .. y0 + TY + HY - 1 (HY = R T) and the 256 columns
// x0 - HX .. x0 - HX + 255 (HX = R T rounded up to a float4). Every level is computed on the
// whole region from the previous level in LDS; level k is exact on the region shrunk by k R
// on each side, so after T levels the output tile is exact, and the values in the shrinking
// rim never reach it. Cells outside the grid stay zero, which is what the unblocked stencil
// reads there. Per cell the arithmetic is the unblocked step's, operation by operation, so
// the levels are bit-identical to T launches of k3_step.
//
// p^{n-1} of a lane's rows is pointwise and stays in registers (so do p^n and m); only the
// level the stencil reads goes through LDS, one buffer with a barrier between the reads of a
// level and the writes of the next. Receiver taps that fall in the output tile are read out of
// LDS at every intermediate level into a tap buffer, which a small kernel turns into trace
// samples with k_sample's tap order; the last level is sampled from memory as usual.
// wavefronts of one blocked workgroup: 16 for short tiles (more wavefronts per SIMD to hide
// LDS latency), 12 for 32-row tiles (4 of the 48 region rows each at depth 4, within the 170
// registers of three wavefronts per SIMD)
template <int TY>
struct BlockWaves {
  static constexpr int waves = TY >= 32 ? 12 : 16;
  static constexpr int threads = 64 * waves;
  // receiver taps one workgroup scans per shot: at least 256 * kTapChunks
  static constexpr int tap_chunks = (kTapChunks * 256 + threads - 1) / threads;
};

template <int TY, int R, int T, int W = BlockWaves<TY>::waves>
struct BlockShape {
  static constexpr int waves = W;
  static constexpr int threads = 64 * W;
  static constexpr int hy = R * T;
  static constexpr int hx = (R * T + 3) / 4 * 4;
  static constexpr int out_w = 256 - 2 * hx;
  static constexpr int rows = TY + 2 * hy;              // region rows
  static constexpr int mr = (rows + waves - 1) / waves;  // rows of one wavefront
  static constexpr int crow = waves * mr;               // computed rows
  static constexpr int lrows = crow + 2 * R;  // LDS rows: R zero rows above and below
};

typedef __attribute__((address_space(1))) void gmem_void;
typedef __attribute__((address_space(3))) void lds_void;

// threadIdx.x as a value the compiler cannot see through: per-thread index arithmetic (the
// lane's column, its receiver-tap slots) is then recomputed where it is used instead of being
// hoisted out of the shot and item loops, where its long live ranges pushed the blocked
// kernels over their register budget (scratch spills).
__device__ __forceinline__ int thread_index() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// Issue the loads of a wavefront's rows c0 .. c0 + mr - 1 of one region of a quiet tile (grid
// rows y .. y + mr - 1, all inside the grid), each lane its own four columns x .. x + 3, straight
// into LDS (global_load_lds: lane l lands 16 bytes after lane l - 1, from the wave-uniform
// address of the row's first output column). The region's margin columns are not loaded: they
// only reach cells within R T of the region's edge. Completion is awaited with s_waitcnt.
template <typename S, int R>
__device__ __forceinline__ void rows_async(const float* P, int y, int x, int P2, int c0,
                                           float* buf) {
#pragma unroll
  for (int k = 0; k < S::mr; ++k) {
    const float* src = P + ((y + k) * P2 + x);
    float* dst = buf + (R + c0 + k) * kLW + 4;
    __builtin_amdgcn_global_load_lds((gmem_void*)src, (lds_void*)dst, 16, 0, 0);
  }
}

// One tile of k4_steps. Q ("quiet") tiles lie inside the grid, margins included, and inside
// the box of neutral damping: every bounds check is dead there and a = b = 1 (taper 1), so
// they step with the arithmetic of the layer-free step, which is bit-identical to the
// damped arithmetic at those coefficients (products by 1 are exact). Quiet tiles also
// double-buffer the region: the next shot's p^n streams into the second LDS buffer and its
// p^{n-1} rows into registers while the current shot computes.
// SP ("split", never quiet): the workgroup steps two half-width tiles of block_cols' split
// column. LDS row r holds region row r of the upper tile in float4 columns 0-32 (lanes 0-31
// and the left margin) and of the lower tile in columns 33-65 (lanes 32-63 and the right
// margin). The x-neighbour reads of lanes 31 and 32 then see the other tile's edge column;
// like the margins, that only reaches region columns within R T of the region's edge, which
// the output columns are not (hx >= R T).
// The tile advances shots [s_first, s_end) (tl.grp is not read); `shots` is the plan's count.
template <int R, int ABS0, int TY, int T, bool Q, bool SP = false>
__device__ __forceinline__ void k4_tile(const Geo<float>& g, const BlockIO& io, int shots,
                                        int s_first, int s_end, const Tile& tl, float* lpa,
                                        float* lpb) {
  static_assert(!(Q && SP), "split tiles lie on the grid's edge");
  constexpr int ABS = Q ? NW_NONE : ABS0;
  using S = BlockShape<TY, R, T>;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int lane = thread_index() & 63, wave = wave_index();
  const int hl = SP ? (lane & 31) : lane;    // lane within its tile
  const int ryA = tl.ry * TY - S::hy;        // grid row of region row 0 (upper tile if SP)
  const int ry0 = SP && lane >= 32 ? ryA + TY : ryA;  // ... of this lane's tile
  const int rx0 = tl.cx * S::out_w - S::hx;  // grid column of region column 0
  const int x = rx0 + 4 * hl;
  const bool inx = Q || (x >= 0 && x < D2);
  const bool outx = inx && hl >= S::hx / 4 && hl < (SP ? 32 : 64) - S::hx / 4;
  const int c0 = wave * S::mr;  // first region row of this wavefront
  const float4 hx = (ABS != NW_NONE && inx) ? ld4(g.damp[2] + x) : f4(0.f);
  const Damp4<ABS> dcol = damp4_col<ABS>(hx);
  // wavelet samples of the block's steps (no vector loads inside the level loop)
  float wl[T];
#pragma unroll
  for (int k = 0; k < T; ++k) {
    const int n = io.n0 + k;
    wl[k] = (io.src_taps > 0 && k < io.levels && n < io.n_src) ? io.wavelet[n] : 0.f;
  }
  float4 mm[S::mr];
#pragma unroll
  for (int k = 0; k < S::mr; ++k) {
    const int y = ry0 + c0 + k;
    mm[k] = (Q || (inx && y >= 0 && y < D1)) ? ld4(io.m + (y * P2 + x)) : f4(0.f);
  }
  auto load_prev = [&](int s, float4* dst) {
    const int64_t base = (int64_t)s * g.cells;
#pragma unroll
    for (int k = 0; k < S::mr; ++k) {
      const int y = ry0 + c0 + k;
      dst[k] = (Q || (inx && y >= 0 && y < D1)) ? ld4(io.p_prev + base + (y * P2 + x))
                                                : f4(0.f);
    }
  };
  float4 pvn[Q ? S::mr : 1];
  if constexpr (Q) {
    rows_async<S, R>(io.p + (int64_t)s_first * g.cells, ry0 + c0, x, P2, c0, lpa);
    load_prev(s_first, pvn);
  }
  // one shot; B selects the LDS buffer at compile time (even shots A, odd shots B), so the
  // compiler sees that the asynchronous fill of the other buffer cannot alias this one's reads
  auto shot = [&](auto bsel, int s) {
    constexpr int B = decltype(bsel)::value;
    const int64_t base = (int64_t)s * g.cells;
    float* lp = (Q && B) ? lpb : lpa;
    float4 pv[S::mr], cur[S::mr];
    if constexpr (Q) {
      __builtin_amdgcn_s_waitcnt(0);  // this shot's region and p^{n-1} rows have landed
      __syncthreads();                // ... for every wavefront; the other buffer is free
#pragma unroll
      for (int k = 0; k < S::mr; ++k) pv[k] = pvn[k];
    } else {
      // the wavefront's own rows of p^n, each lane its four columns (zero outside the grid);
      // they are also the lane's current level
      const float* P = io.p + base;
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const int y = ry0 + c0 + k;
        cur[k] = (inx && y >= 0 && y < D1) ? ld4(P + (y * P2 + x)) : f4(0.f);
      }
      load_prev(s, pv);
      __syncthreads();  // the previous shot's reads of lp are done
#pragma unroll
      for (int k = 0; k < S::mr; ++k)
        *reinterpret_cast<float4*>(lp + (R + c0 + k) * kLW + 4 * lane + 4) = cur[k];
    }
    // does a source tap of this shot fall in this wavefront's rows of the region? (uniform)
    bool src_rows = false;
    for (int t = 0; t < io.src_taps; ++t) {
      const int idx = ldc(io.src_idx, s * io.src_taps + t);
      const int sy = idx / P2 - ry0 - c0;
      src_rows |= sy >= 0 && sy < S::mr;
    }
    // receiver taps of this shot inside the output tile: LDS offset, or -1
    int toff[BlockWaves<TY>::tap_chunks];
    int n_tap = io.tapbuf ? io.rec_count[s] * io.taps : 0;
    if (n_tap > 0) {
      // no scan when the shot's taps all lie outside the tile's output rows (uniform)
      const int o0 = ryA + S::hy, o1 = o0 + (SP ? 2 * TY : TY);
      if (ldc(io.tap_rows, 2 * s + 1) < o0 || ldc(io.tap_rows, 2 * s) >= o1) n_tap = 0;
    }
    const int tap0 = io.tapbuf ? io.rec_first[s] * io.taps : 0;
#pragma unroll
    for (int c = 0; c < BlockWaves<TY>::tap_chunks; ++c) {
      const int e = thread_index() + S::threads * c;
      toff[c] = -1;
      if (e < n_tap) {
        const int idx = io.rec_idx[tap0 + e];
        const int ty = idx / P2, tx = idx - ty * P2;
        if constexpr (SP) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int ryh = ryA + h * TY;
            if (ty >= ryh + S::hy && ty < ryh + S::hy + TY && tx >= rx0 + S::hx &&
                tx < rx0 + 128 - S::hx)
              toff[c] = (R + ty - ryh) * kLW + 4 + 128 * h + (tx - rx0);
          }
        } else if (ty >= ry0 + S::hy && ty < ry0 + S::hy + TY && tx >= rx0 + S::hx &&
                   tx < rx0 + 256 - S::hx) {
          toff[c] = (R + ty - ry0) * kLW + 4 + (tx - rx0);
        }
      }
    }
    if constexpr (!Q) __syncthreads();
    if constexpr (Q) {
      if (s + 1 < s_end) {  // stream the next shot in while this one computes
        rows_async<S, R>(io.p + base + g.cells, ry0 + c0, x, P2, c0, B ? lpa : lpb);
        load_prev(s + 1, pvn);
      }
#pragma unroll
      for (int k = 0; k < S::mr; ++k) cur[k] = lrow(lp, R + c0 + k, lane);
    }
    if (io.ck_prev) {
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const int c = c0 + k, y = ry0 + c;
        if (outx && c >= S::hy && c < S::hy + TY && (Q || y < D1)) {
          st4(io.ck_prev + base + (y * P2 + x), pv[k]);
          st4(io.ck_cur + base + (y * P2 + x), cur[k]);
        }
      }
    }
    for (int lvl = 1; lvl <= io.levels; ++lvl) {
      const int n = io.n0 + lvl - 1;
      const bool src = src_rows && n < io.n_src;
      float w_n = 0.f;
#pragma unroll
      for (int k = 0; k < T; ++k) w_n = (k == lvl - 1) ? wl[k] : w_n;
      const int tpv = n > io.first_step;
      float* lap_out = io.lap_out[lvl];
      float4 w[2 * R + 1];
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) w[j] = lrow(lp, c0 + j, lane);
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const int y = ry0 + c0 + k;
        w[2 * R] = lrow(lp, c0 + k + 2 * R, lane);
        float4 o = f4(0.f);
        // rows that can still reach the output tile: [R lvl, rows - R lvl) (uniform); the rim
        // outside keeps stale values, which no later level reads
        const bool live = c0 + k >= R * lvl && c0 + k < S::rows - R * lvl;
        if (live && (Q || (y >= 0 && y < D1))) {  // columns outside the grid are zeroed below
          const float4 nb = lnb(lp, c0 + k + R, lane);
          const float4 lap = stencil4<R>(g, w, nb.x, nb.y, nb.z, nb.w);
          const int c = c0 + k;
          if (lap_out && outx && c >= S::hy && c < S::hy + TY && (Q || y < D1))
            st4(lap_out + base + (y * P2 + x), lap);
          const float hy = ABS != NW_NONE ? (SP ? g.damp[1][y] : ldc(g.damp[1], y)) : 0.f;
          const Damp4<ABS> dm = damp4_row<ABS>(hy, hx, dcol);
          const float4 pc = cur[k];
          o = make_float4(rhs1<ABS>(pc.x, pv[k].x, mm[k].x, lap.x, dm.b.x, dm.tap.x, tpv),
                          rhs1<ABS>(pc.y, pv[k].y, mm[k].y, lap.y, dm.b.y, dm.tap.y, tpv),
                          rhs1<ABS>(pc.z, pv[k].z, mm[k].z, lap.z, dm.b.z, dm.tap.z, tpv),
                          rhs1<ABS>(pc.w, pv[k].w, mm[k].w, lap.w, dm.b.w, dm.tap.w, tpv));
          if (src) {
            const int ro = y * P2 + x;
            for (int t = 0; t < io.src_taps; ++t) {
              const int d = ldc(io.src_idx, s * io.src_taps + t) - ro;
              if (d >= 0 && d < 4) {
                const float sv = ldc(io.src_val, s * io.src_taps + t) * w_n;
                o.x += d == 0 ? sv : 0.f;
                o.y += d == 1 ? sv : 0.f;
                o.z += d == 2 ? sv : 0.f;
                o.w += d == 3 ? sv : 0.f;
              }
            }
          }
          if (ABS != NW_NONE) o = mul4(dm.a, o);
          if (!Q && !inx) o = f4(0.f);
        }
        pv[k] = o;  // p^{n-1} is consumed: the slot takes the new level
#pragma unroll
        for (int j = 0; j < 2 * R; ++j) w[j] = w[j + 1];
      }
      if (lvl < io.levels) {
        lds_barrier();  // every read of the previous level is done
#pragma unroll
        for (int k = 0; k < S::mr; ++k)
          *reinterpret_cast<float4*>(lp + (R + c0 + k) * kLW + 4 * lane + 4) = pv[k];
        lds_barrier();
        if (io.tapbuf) {
          float* tb = io.tapbuf + ((int64_t)(lvl - 1) * shots + s) * io.tap_stride;
#pragma unroll
          for (int c = 0; c < BlockWaves<TY>::tap_chunks; ++c)
            if (toff[c] >= 0) tb[thread_index() + S::threads * c] = lp[toff[c]];
        }
      }
      float* out = io.out[lvl];
      if (out) {
#pragma unroll
        for (int k = 0; k < S::mr; ++k) {
          const int c = c0 + k, y = ry0 + c;
          if (outx && c >= S::hy && c < S::hy + TY && (Q || y < D1))
            st4(out + base + (y * P2 + x), pv[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const float4 t = cur[k];
        cur[k] = pv[k];
        pv[k] = t;
      }
    }
  };
  for (int s = s_first; s < s_end; s += 2) {
    shot(std::integral_constant<int, 0>{}, s);
    if (s + 1 < s_end) shot(std::integral_constant<int, 1>{}, s + 1);
  }
}

// The R zero rows above and below the region and the 4 margin columns on each side of every
// row in both LDS buffers (the stencil of the region's edge rows and columns reads them); tiles
// never write them.
template <class S, int R>
__device__ __forceinline__ void zero_margin_rows(float* lpa, float* lpb) {
  for (int i = threadIdx.x; i < R * kLW; i += S::threads) {
    lpa[i] = 0.f;
    lpa[(S::lrows - R) * kLW + i] = 0.f;
    lpb[i] = 0.f;
    lpb[(S::lrows - R) * kLW + i] = 0.f;
  }
  for (int r = R + threadIdx.x; r < S::lrows - R; r += S::threads) {
    st4_lds(lpa + r * kLW, f4(0.f));
    st4_lds(lpa + r * kLW + kLW - 4, f4(0.f));
    st4_lds(lpb + r * kLW, f4(0.f));
    st4_lds(lpb + r * kLW + kLW - 4, f4(0.f));
  }
}

// Shots [s_first, s_end) of tile tl, with the variant its position selects: split column,
// quiet (region inside the grid and the box of neutral damping) or boundary.
template <int R, int ABS, int TY, int T>
__device__ __forceinline__ void k4_any_tile(const Geo<float>& g, const BlockIO& io, int shots,
                                            int s_first, int s_end, const Tile& tl,
                                            const BlockCols& cols, float* lpa, float* lpb) {
  using S = BlockShape<TY, R, T>;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2];
  const int ry0 = tl.ry * TY - S::hy, rx0 = tl.cx * S::out_w - S::hx;
  const bool quiet = ry0 >= 0 && ry0 + S::crow <= D1 && rx0 >= 4 && rx0 + 260 <= D2 &&
                     ry0 >= g.quiet[1][0] && ry0 + S::crow <= g.quiet[1][1] &&
                     rx0 >= g.quiet[2][0] && rx0 + 256 <= g.quiet[2][1];
  if (cols.split && tl.cx == cols.full)
    k4_tile<R, ABS, TY, T, false, true>(g, io, shots, s_first, s_end, tl, lpa, lpb);
  else if (quiet)
    k4_tile<R, ABS, TY, T, true>(g, io, shots, s_first, s_end, tl, lpa, lpb);
  else
    k4_tile<R, ABS, TY, T, false>(g, io, shots, s_first, s_end, tl, lpa, lpb);
}

template <int R, int ABS, int TY, int T>
__global__ void __launch_bounds__(BlockWaves<TY>::threads) k4_steps(Geo<float> g, BlockIO io, int shots,
                                                          int spw, int xcd_map) {
  using S = BlockShape<TY, R, T>;
  __shared__ __align__(16) float lpa[S::lrows * kLW];
  __shared__ __align__(16) float lpb[S::lrows * kLW];
  const BlockCols cols = block_cols<S>((int)g.D[2]);
  const Tile tl = tile_of_cols(cols, div_up((int)g.D[1], TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;  // uniform over the workgroup
  zero_margin_rows<S, R>(lpa, lpb);
  const int s_first = tl.grp * spw;
  k4_any_tile<R, ABS, TY, T>(g, io, shots, s_first, min(shots, s_first + spw), tl, cols, lpa,
                             lpb);
}

// k4_steps as a work queue. One launch has as many workgroups as fit on the device at once;
// each takes (tile, shot range) items until every queue is empty, so no workgroup waits for
// another and the device is not idle between tiles while the dispatcher starts the next
// workgroup (about 3 us per tile on MI355X), nor in a tail of a few long tiles. Queue j
// holds the tiles k4_steps's XCD map gives XCD j, in the same order (row strips that share
// halo rows run on one XCD and share its L2), the last ones split into single shots. A
// workgroup serves the queue of the XCD it runs on (hardware register XCC_ID), then takes
// from the others in turn. The items' cells, arithmetic and outputs are those of k4_steps.
// Counters: ctr[j] is the next item of queue j, ctr[kXcds] counts the workgroups that found
// every queue empty; the last of them sets all back to zero for the next launch, so launches
// that share ctr must not overlap (one plan's launches are stream ordered).
struct BlockQueues {
  int off[kXcds + 1];  // items of queue j: [off[j], off[j + 1])
};
static_assert(kQueues == kXcds, "one work queue per XCD");

template <int R, int ABS, int TY, int T>
__global__ void __launch_bounds__(BlockWaves<TY>::threads)
    k4_queue(Geo<float> g, BlockIO io, int shots, const int4* __restrict__ items,
             BlockQueues q, int* ctr) {
  using S = BlockShape<TY, R, T>;
  __shared__ __align__(16) float lpa[S::lrows * kLW];
  __shared__ __align__(16) float lpb[S::lrows * kLW];
  __shared__ int item;
  const BlockCols cols = block_cols<S>((int)g.D[2]);
  zero_margin_rows<S, R>(lpa, lpb);
  unsigned xcc;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
  int tried = 0;  // queues found empty (thread 0)
  for (;;) {
    __syncthreads();  // every wavefront is done with the previous item and its LDS reads
    if (threadIdx.x == 0) {
      int it = -1;
      while (tried < kXcds) {
        // unsigned modulo: j stays in [0, kXcds) whatever the register's upper bits hold
        const int j = (int)((xcc + (unsigned)tried) % (unsigned)kXcds);
        const int k = atomicAdd(&ctr[j], 1);
        if (k < q.off[j + 1] - q.off[j]) {
          it = q.off[j] + k;
          break;
        }
        ++tried;
      }
      item = it;
    }
    __syncthreads();
    const int it = item;
    if (it < 0) break;  // uniform: every queue is empty
    const int4 w = items[it];  // row strip, column chunk, first shot, end shot
    Tile tl;
    tl.ry = w.x;
    tl.cx = w.y;
    tl.grp = 0;
    tl.valid = true;
    k4_any_tile<R, ABS, TY, T>(g, io, shots, w.z, w.w, tl, cols, lpa, lpb);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&ctr[kXcds], 1) == (int)gridDim.x - 1) {  // the last workgroup out
      for (int j = 0; j <= kXcds; ++j) ctr[j] = 0;
      __threadfence();
    }
  }
}

// ------------------------------------------------------------------------------------------
// temporally blocked reverse steps
// ------------------------------------------------------------------------------------------
// The reverse step has the forward step's shape: lambda^n from the stencil of
// q = (m a) lambda^{n+1} and pointwise lambda^{n+1}, lambda^{n+2}. k5_tile advances a tile
// through up to T reverse steps the way k4_tile advances p (same region, halo and LDS
// exchange of the stencil operand; q instead of p), with two additions per level:
// * the imaging sum a lambda^{n+1} * (Laplacian of p^n) over the output cells, from the
//   Laplacians the forward sweep kept (pointwise), accumulated in float64 registers over the
//   levels and the shot group and added to the group's gradient slice once;
// * the residual injection of a recorded step: the level's lambda goes through LDS, the
//   workgroup's threads add the precomputed sums at the injection cells inside the region
//   (halo included, so the redundant cells stay exact), and the owners read it back.
// The lambda values are bit-identical to single reverse steps followed by k_inject.
// wavefronts of a blocked reverse workgroup: 12 for 16-row tiles (two region rows each at depth
// 2, within the registers of three wavefronts per SIMD), 8 for 8-row tiles
template <int TY>
struct AdjBlockWaves {
  static constexpr int waves = TY >= 16 ? 12 : 8;
};

template <int R, int ABS0, int TY, int T, bool Q>
__device__ __forceinline__ void k5_tile(const Geo<float>& g, const AdjBlockIO& io, int shots,
                                        int spw, const Tile& tl, float* lpa, float* lpb) {
  constexpr int ABS = Q ? NW_NONE : ABS0;
  using S = BlockShape<TY, R, T, AdjBlockWaves<TY>::waves>;
  constexpr int kInjChunks = (kTapChunks * 256 + S::threads - 1) / S::threads;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int ry0 = tl.ry * TY - S::hy;
  const int rx0 = tl.cx * S::out_w - S::hx;
  const int x = rx0 + 4 * lane;
  const bool inx = Q || (x >= 0 && x < D2);
  const bool outx = inx && lane >= S::hx / 4 && lane < 64 - S::hx / 4;
  const int c0 = wave * S::mr;
  const float4 hx = (ABS != NW_NONE && inx) ? ld4(g.damp[2] + x) : f4(0.f);
  const Damp4<ABS> dcol = damp4_col<ABS>(hx);
  float4 mm[S::mr];  // m a of this wavefront's rows (shot-independent)
#pragma unroll
  for (int k = 0; k < S::mr; ++k) {
    const int y = ry0 + c0 + k;
    mm[k] = (Q || (inx && y >= 0 && y < D1)) ? ld4(io.ma + (y * P2 + x)) : f4(0.f);
  }
  double2 ga[S::mr], gb[S::mr];  // float64 imaging sums over levels and shots
#pragma unroll
  for (int k = 0; k < S::mr; ++k) {
    ga[k] = make_double2(0.0, 0.0);
    gb[k] = ga[k];
  }
  auto load_lam2 = [&](int s, float4* dst) {
    const int64_t base = (int64_t)s * g.cells;
#pragma unroll
    for (int k = 0; k < S::mr; ++k) {
      const int y = ry0 + c0 + k;
      dst[k] = (Q || (inx && y >= 0 && y < D1)) ? ld4nt(io.lam2 + base + (y * P2 + x)) : f4(0.f);
    }
  };
  // kept Laplacians of every level for the wavefront's output rows, loaded with the shot's
  // region so the level loop does not wait on memory (zero where no imaging applies)
  auto load_laps = [&](int s, float4 (*dst)[S::mr]) {
    const int64_t base = (int64_t)s * g.cells;
#pragma unroll
    for (int lv = 0; lv < T; ++lv) {
      const float* lapl = io.lap[lv + 1];
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const int c = c0 + k, y = ry0 + c;
        const bool ok = lv < io.levels && lapl && outx && c >= S::hy && c < S::hy + TY &&
                        (Q || y < D1);
        dst[lv][k] = ok ? ld4(lapl + base + (y * P2 + x)) : f4(0.f);
      }
    }
  };
  const int s_first = tl.grp * spw;
  const int s_end = min(shots, s_first + spw);
  float4 pvn[Q ? S::mr : 1];
  float4 lapn[Q ? T : 1][S::mr];
  if constexpr (Q) {
    rows_async<S, R>(io.lam1 + (int64_t)s_first * g.cells, ry0 + c0, x, P2, c0, lpa);
    load_lam2(s_first, pvn);
    load_laps(s_first, lapn);
  }
  auto shot = [&](auto bsel, int s) {
    constexpr int B = decltype(bsel)::value;
    const int64_t base = (int64_t)s * g.cells;
    float* lp = (Q && B) ? lpb : lpa;
    float4 pv[S::mr], cur[S::mr];  // lambda^{n+2}, lambda^{n+1} of the current level
    float4 lapv[T][S::mr];         // kept Laplacians of the levels' output rows
    if constexpr (Q) {
      __builtin_amdgcn_s_waitcnt(0);  // this shot's region, rows and Laplacians have landed
      __syncthreads();
#pragma unroll
      for (int k = 0; k < S::mr; ++k) pv[k] = pvn[k];
#pragma unroll
      for (int lv = 0; lv < T; ++lv)
#pragma unroll
        for (int k = 0; k < S::mr; ++k) lapv[lv][k] = lapn[lv][k];
    } else {
      // the wavefront's own region rows of lambda^{nh}, each lane its four columns (zero
      // outside the grid); they are also the lane's current level
      const float* L1 = io.lam1 + base;
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const int y = ry0 + c0 + k;
        cur[k] = (inx && y >= 0 && y < D1) ? ld4(L1 + (y * P2 + x)) : f4(0.f);
      }
      load_lam2(s, pv);
      load_laps(s, lapv);
      __syncthreads();  // the previous shot's reads of lp are done
#pragma unroll
      for (int k = 0; k < S::mr; ++k)
        *reinterpret_cast<float4*>(lp + (R + c0 + k) * kLW + 4 * lane + 4) = cur[k];
    }
    // injection cells of this shot inside the region: LDS offset (or -1) and value index
    int ioff[kInjChunks];
    bool any_inj = false;
    int n_inj = io.injv ? io.inj_count[s] : 0;
    // no scan when the shot's injection cells all lie outside the region's rows (uniform)
    if (n_inj > 0 && (ldc(io.inj_rows, 2 * s + 1) < ry0 || ldc(io.inj_rows, 2 * s) >= ry0 + S::crow))
      n_inj = 0;
    const int inj0 = io.injv ? io.inj_first[s] : 0;
#pragma unroll
    for (int c = 0; c < kInjChunks; ++c) {
      const int u = threadIdx.x + S::threads * c;
      ioff[c] = -1;
      if (u < n_inj) {
        const int idx = io.inj_cell[inj0 + u];
        const int ty = idx / P2, tx = idx - ty * P2;
        if (ty >= ry0 && ty < ry0 + S::crow && tx >= rx0 && tx < rx0 + 256) {
          ioff[c] = (R + ty - ry0) * kLW + 4 + (tx - rx0);
          any_inj = true;
        }
      }
    }
    if constexpr (!Q) __syncthreads();
    if constexpr (Q) {
      if (s + 1 < s_end) {  // stream the next shot in while this one computes
        rows_async<S, R>(io.lam1 + base + g.cells, ry0 + c0, x, P2, c0, B ? lpa : lpb);
        load_lam2(s + 1, pvn);
        load_laps(s + 1, lapn);
      }
      // lambda^{nh} of this wavefront's rows
#pragma unroll
      for (int k = 0; k < S::mr; ++k) cur[k] = lrow(lp, R + c0 + k, lane);
    }
    // the stencil operand q = (m a) lambda in place
    const bool inj_tile = __syncthreads_or(any_inj);  // (also: every read of lambda is done)
#pragma unroll
    for (int k = 0; k < S::mr; ++k)
      *reinterpret_cast<float4*>(lp + (R + c0 + k) * kLW + 4 * lane + 4) = mul4(mm[k], cur[k]);
    lds_barrier();
#pragma unroll
    for (int lvl = 1; lvl <= T; ++lvl) {
      if (lvl > io.levels) break;  // uniform
      const int nk = io.nh - lvl;
      const int tpn = (nk + 1) > io.first_step;
      const float* lapl = io.lap[lvl];
      float4 w[2 * R + 1];
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) w[j] = lrow(lp, c0 + j, lane);
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const int y = ry0 + c0 + k;
        w[2 * R] = lrow(lp, c0 + k + 2 * R, lane);
        float4 o = f4(0.f);
        // rows that can still reach the output tile: [R lvl, rows - R lvl) (uniform); the rim
        // outside takes zeros, which no later level reads into an output row
        const bool live = c0 + k >= R * lvl && c0 + k < S::rows - R * lvl;
        if (live && (Q || (y >= 0 && y < D1))) {
          const float4 nb = lnb(lp, c0 + k + R, lane);
          const float4 lapq = stencil4<R>(g, w, nb.x, nb.y, nb.z, nb.w);
          const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
          const Damp4<ABS> dm = damp4_row<ABS>(hy, hx, dcol);
          const float4 lc = cur[k], l2c = pv[k];
          float4 rb;
          o = make_float4(adj1<ABS>(lc.x, l2c.x, lapq.x, dm.a.x, dm.b.x, dm.tap.x, tpn, rb.x),
                          adj1<ABS>(lc.y, l2c.y, lapq.y, dm.a.y, dm.b.y, dm.tap.y, tpn, rb.y),
                          adj1<ABS>(lc.z, l2c.z, lapq.z, dm.a.z, dm.b.z, dm.tap.z, tpn, rb.z),
                          adj1<ABS>(lc.w, l2c.w, lapq.w, dm.a.w, dm.b.w, dm.tap.w, tpn, rb.w));
          if (!Q && !inx) o = f4(0.f);
          const int c = c0 + k;
          if (lapl && outx && c >= S::hy && c < S::hy + TY && (Q || y < D1)) {
            const float4 lapp = lapv[lvl - 1][k];
            ga[k].x = __builtin_fma((double)rb.x, (double)lapp.x, ga[k].x);
            ga[k].y = __builtin_fma((double)rb.y, (double)lapp.y, ga[k].y);
            gb[k].x = __builtin_fma((double)rb.z, (double)lapp.z, gb[k].x);
            gb[k].y = __builtin_fma((double)rb.w, (double)lapp.w, gb[k].y);
          }
        }
        pv[k] = o;  // lambda^{n+2} is consumed: the slot takes lambda^n
#pragma unroll
        for (int j = 0; j < 2 * R; ++j) w[j] = w[j + 1];
      }
      const bool inject = inj_tile && io.inj_rec[lvl];
      if (inject || lvl < io.levels) lds_barrier();  // every read of q^{n+1} is done
      if (inject) {  // lambda^n += the residual sums, through LDS
#pragma unroll
        for (int k = 0; k < S::mr; ++k)
          *reinterpret_cast<float4*>(lp + (R + c0 + k) * kLW + 4 * lane + 4) = pv[k];
        lds_barrier();
        const float* iv = io.injv + ((int64_t)(lvl - 1) * shots + s) * io.inj_max;
#pragma unroll
        for (int c = 0; c < kInjChunks; ++c)
          if (ioff[c] >= 0) lp[ioff[c]] = lp[ioff[c]] + iv[threadIdx.x + S::threads * c];
        lds_barrier();
#pragma unroll
        for (int k = 0; k < S::mr; ++k) pv[k] = lrow(lp, R + c0 + k, lane);
        lds_barrier();
      }
      if (lvl < io.levels) {
#pragma unroll
        for (int k = 0; k < S::mr; ++k)
          *reinterpret_cast<float4*>(lp + (R + c0 + k) * kLW + 4 * lane + 4) =
              mul4(mm[k], pv[k]);
        lds_barrier();
      }
      float* out = io.out[lvl];
      if (out) {
#pragma unroll
        for (int k = 0; k < S::mr; ++k) {
          const int c = c0 + k, y = ry0 + c;
          if (outx && c >= S::hy && c < S::hy + TY && (Q || y < D1))
            st4(out + base + (y * P2 + x), pv[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < S::mr; ++k) {
        const float4 t = cur[k];
        cur[k] = pv[k];
        pv[k] = t;
      }
    }
  };
  for (int s = s_first; s < s_end; s += 2) {
    shot(std::integral_constant<int, 0>{}, s);
    if (s + 1 < s_end) shot(std::integral_constant<int, 1>{}, s + 1);
  }
  if (io.grad) {
    double* G = io.grad + (int64_t)tl.grp * g.cells;
#pragma unroll
    for (int k = 0; k < S::mr; ++k) {
      const int c = c0 + k, y = ry0 + c;
      if (outx && c >= S::hy && c < S::hy + TY && (Q || y < D1)) {
        double2* gp = reinterpret_cast<double2*>(G + (y * P2 + x));
        double2 a = gp[0], b = gp[1];
        a.x += ga[k].x;
        a.y += ga[k].y;
        b.x += gb[k].x;
        b.y += gb[k].y;
        gp[0] = a;
        gp[1] = b;
      }
    }
  }
}

template <int R, int ABS, int TY, int T>
__global__ void __launch_bounds__(64 * AdjBlockWaves<TY>::waves) k5_adjoint(Geo<float> g, AdjBlockIO io,
                                                             int shots, int spw, int xcd_map) {
  using S = BlockShape<TY, R, T, AdjBlockWaves<TY>::waves>;
  __shared__ __align__(16) float lpa[S::lrows * kLW];
  __shared__ __align__(16) float lpb[S::lrows * kLW];
  const int D1 = (int)g.D[1], D2 = (int)g.D[2];
  const Tile tl = tile_of(div_up(D2, S::out_w), div_up(D1, TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;  // uniform over the workgroup
  zero_margin_rows<S, R>(lpa, lpb);
  const int ry0 = tl.ry * TY - S::hy, rx0 = tl.cx * S::out_w - S::hx;
  const bool quiet = ry0 >= 0 && ry0 + S::crow <= D1 && rx0 >= 4 && rx0 + 260 <= D2 &&
                     ry0 >= g.quiet[1][0] && ry0 + S::crow <= g.quiet[1][1] &&
                     rx0 >= g.quiet[2][0] && rx0 + 256 <= g.quiet[2][1];
  if (quiet)
    k5_tile<R, ABS, TY, T, true>(g, io, shots, spw, tl, lpa, lpb);
  else
    k5_tile<R, ABS, TY, T, false>(g, io, shots, spw, tl, lpa, lpb);
}

// trace samples of the intermediate levels of a blocked launch, from its tap buffer: the
// products and the tap sum of k_sample, in its order
__global__ void __launch_bounds__(256) k_sample_levels(Geo<float> g, SampleIO<float> io,
                                                       const float* tapbuf, int tap_stride,
                                                       int shots, LevelRecords lj) {
  (void)g;
  const int s = blockIdx.y, lvl = blockIdx.z;
  const int j = lj.j[lvl];
  if (j < 0) return;  // uniform over the workgroup
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int taps = io.taps;
  const int r = t / taps, k = t % taps;
  const int count = io.rec_count[s];
  float v = 0.f;
  if (r < count) {
    const int64_t row = (int64_t)io.rec_first[s] + r;
    v = io.rec_w[row * taps + k] * tapbuf[((int64_t)lvl * shots + s) * tap_stride + t];
  }
  for (int off = 1; off < taps; off <<= 1) v += __shfl_xor(v, off, 64);
  if (r < count && k == 0) io.traces[((int64_t)io.trace_first[s] + r) * io.n_t + j] = v;
}

// ------------------------------------------------------------------------------------------
// PML (2-D): LDS-tiled forward and reverse steps with the face fields
// ------------------------------------------------------------------------------------------
// Face fields psi_a live on the faces normal to axis a: psi_y rows of D2 faces (D1 + 1 rows),
// psi_x rows of D2 + 1 faces (D1 rows) stored FP2 apart (a multiple of 4), so a lane's faces
// x..x+3 are one aligned float4. A lane's four cells x..x+3 of row y touch x-faces x..x+4 (face
// x+4 is the next lane's face x: it takes it from that lane with a shuffle and computes it
// again, with the same operations; lane 63 and the grid's last lane load it) and y-faces of
// rows y and y+1 (the next row of the wavefront computes row y+1 again). Each face is written by the cell on its high side; the last cell of a row (column)
// also writes the final face. Per face f between cells f-1 and f:
//   forward:  psi'_f = alpha_f psi_f + (zeta_other - zf_f) kappa_f (p_f - p_{f-1})
//             lap += ((psi'_{f+1} + psi_{f+1}) - (psi'_f + psi_f)) / (2h)   (per cell)
//   reverse:  t_f = psibar_f - (q_f - q_{f-1}) / (2h)
//             lambda -= (zeta_other - zf_{f+1}) kappa_{f+1} t_{f+1} - (...)_f t_f
//             psibar'_f = alpha_f t_f - (q_f - q_{f-1}) / (2h)
// following nefi/physics/wave.py:529-552 and its adjoint, in the generic kernels' order
// (Laplacian, then the y faces, then the x faces).
struct PmlX {  // x-face profiles of a lane's faces x..x+4
  float a[5], k[5], zf[5];
};

__device__ __forceinline__ PmlX pml_x(const Geo<float>& g, int x, bool act) {
  PmlX r;
  const int D2 = (int)g.D[2];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const bool ok = act && x + j <= D2;
    r.a[j] = ok ? g.alpha[2][x + j] : 0.f;
    r.k[j] = ok ? g.kappa[2][x + j] : 0.f;
    r.zf[j] = ok ? g.zf[2][x + j] : 0.f;
  }
  return r;
}

// Float4 tile loads of a workgroup of S::threads lanes: rows y0 - R .. y0 + TY + R - 1 and
// columns x0 - 4 .. x0 + 259 of one grid (zero outside it), element i of the LDS tile per lane
// i, i + threads, ...
template <typename S, int R>
__device__ __forceinline__ void tile_load_w(const float* src, int y0, int x0, int D1, int D2,
                                            int P2, float4* v) {
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + S::threads * k;
    const int row = i / (kLW / 4), c4 = i % (kLW / 4);
    const int gy = y0 - R + row, gx = x0 - 4 + 4 * c4;
    const bool ok = i < S::n4 && gy >= 0 && gy < D1 && gx >= 0 && gx < D2;
    v[k] = ok ? ld4(src + (gy * P2 + gx)) : f4(0.f);
  }
}

template <typename S>
__device__ __forceinline__ void tile_store_w(float* lds, const float4* v) {
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + S::threads * k;
    if (i < S::n4) *reinterpret_cast<float4*>(lds + 4 * i) = v[k];
  }
}

// The x-faces x..x+4 of a lane from its aligned float4 of faces x..x+3 (lanes whose four
// cells lie in the quiet box pass zeros, the faces' values there) and face x+4 taken from the
// next lane, or from `edge` (loaded by lane 63 and the grid's last lane).
__device__ __forceinline__ void faces5(const float4& v, float edge, bool own_edge, float* f) {
  const float nx = __shfl_down(v.x, 1, 64);
  f[0] = v.x;
  f[1] = v.y;
  f[2] = v.z;
  f[3] = v.w;
  f[4] = own_edge ? edge : nx;
}

// One tile of k3_step_pml. Q: the tile lies in the PML's quiet box, where every face-field
// coefficient and the damping vanish: it neither reads nor writes face fields (they are zero
// there and stay zero; face buffers, snapshots and keep slots hold zeros in the box from
// allocation) and steps with a = b = 1, which are the values the damped arithmetic takes there.
template <int R, int TY, bool Q>
__device__ __forceinline__ void k3_step_pml_tile(const Geo<float>& g, const StepIO<float>& io,
                                                 int shots, int spw, const Tile& tl, float* lp) {
  using S = TileShape<TY, R>;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int F2 = (int)g.FP2;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  const int x = x0 + 4 * lane;
  const int r0 = wave * S::my_rows;
  const bool act = x < D2;
  const bool last_col = x + 4 == D2;
  // lanes that load their face x+4 (the next lane's face x lies in another tile or past D2)
  const bool own_edge = lane == 63 || last_col;
  const float4 hx = act ? ld4(g.damp[2] + x) : f4(0.f);
  const Damp4<NW_PML> dcol = damp4_col<NW_PML>(hx);
  const float4 zcx = act ? ld4(g.zc[2] + x) : f4(0.f);
  const PmlX px = pml_x(g, x, act);
  const float ih1 = g.inv2h[1], ih2 = g.inv2h[2];
  const float w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.f;
  // lanes whose four cells lie in the quiet box (within a tile that does not) skip their face
  // loads and stores the same way; a row's cells are in the box when the row is
  const bool col_quiet = x >= g.quiet[2][0] && x + 4 <= g.quiet[2][1];
  float4 mm[S::my_rows];
#pragma unroll
  for (int k = 0; k < S::my_rows; ++k) {
    const int y = y0 + r0 + k;
    mm[k] = (act && y < D1) ? ld4(io.m + (y * P2 + x)) : f4(0.f);
  }
  const int s_end = min(shots, (tl.grp + 1) * spw);
  for (int s = tl.grp * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const int64_t fy = (int64_t)s * g.fcells[1], fx = (int64_t)s * g.fcells[2];
    float4 v[S::per_thread];
    tile_load<TY, R>(io.p + base, y0, x0, D1, D2, P2, v);
    float4 pp[S::my_rows];
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      pp[k] = (act && y < D1) ? ld4nt(io.p_prev + base + (y * P2 + x)) : f4(0.f);
    }
    // The face fields the wave's rows read, issued with the tile so that a shot costs one
    // memory round trip: y-face rows y0+r0 .. y0+r0+my_rows (cell row k reads face rows k and
    // k + 1) and each cell row's x-faces. A row whose lane lies in the quiet box reads no faces;
    // faces skipped or read for such rows hold zero there.
    float4 oyv[Q ? 1 : S::my_rows + 1], oxv[Q ? 1 : S::my_rows];
    float oxe[Q ? 1 : S::my_rows];
    if constexpr (!Q) {
      auto row_faces = [&](int k) {
        const int y = y0 + r0 + k;
        return act && y < D1 && !(col_quiet && y >= g.quiet[1][0] && y < g.quiet[1][1]);
      };
#pragma unroll
      for (int j = 0; j <= S::my_rows; ++j) {
        const int y = y0 + r0 + j;
        const bool used = (j < S::my_rows && row_faces(j)) || (j > 0 && row_faces(j - 1));
        oyv[j] = used ? ld4(io.psi_old[1] + fy + ((int64_t)y * D2 + x)) : f4(0.f);
      }
#pragma unroll
      for (int k = 0; k < S::my_rows; ++k) {
        const float* ox_row = io.psi_old[2] + fx + (int64_t)(y0 + r0 + k) * F2 + x;
        const bool faces = row_faces(k);
        oxv[k] = faces ? ld4nt(ox_row) : f4(0.f);  // read once: streaming
        oxe[k] = (faces && own_edge && x + 4 <= D2) ? ox_row[4] : 0.f;
      }
    }
    lds_barrier();  // the previous shot's reads of lp are done
    tile_store<TY, R>(lp, v);
    lds_barrier();
    float4 w[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) w[k] = lrow(lp, r0 + k, lane);
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      w[2 * R] = lrow(lp, r0 + k + 2 * R, lane);
      if (act && y < D1) {
        const float4 nb = lnb(lp, r0 + k + R, lane);
        float4 lap = stencil4<R>(g, w, nb.x, nb.y, nb.z, nb.w);
        const float4 pc = w[R], pm = w[R - 1], pq = w[R + 1];
        const float zcy = Q ? 0.f : ldc(g.zc[1], y);
        if constexpr (!Q) {
          const bool lq = col_quiet && y >= g.quiet[1][0] && y < g.quiet[1][1];
          // y faces of rows y (own) and y + 1
          const float ay0 = ldc(g.alpha[1], y), ay1 = ldc(g.alpha[1], y + 1);
          const float ky0 = ldc(g.kappa[1], y), ky1 = ldc(g.kappa[1], y + 1);
          const float zy0 = ldc(g.zf[1], y), zy1 = ldc(g.zf[1], y + 1);
          const float4 oy0 = oyv[k], oy1 = oyv[k + 1];
          float4 ny0, ny1;
#define NW_PML_Y(c)                                                                          \
  {                                                                                          \
    ny0.c = __builtin_fmaf((zcx.c - zy0) * ky0, pc.c - pm.c, ay0 * oy0.c);                   \
    ny1.c = __builtin_fmaf((zcx.c - zy1) * ky1, pq.c - pc.c, ay1 * oy1.c);                   \
    lap.c = lap.c + ((ny1.c + oy1.c) - (ny0.c + oy0.c)) * ih1;                               \
  }
          NW_PML_Y(x) NW_PML_Y(y) NW_PML_Y(z) NW_PML_Y(w)
#undef NW_PML_Y
          if (!lq) {
            st4(io.psi_new[1] + fy + ((int64_t)y * D2 + x), ny0);
            if (y == D1 - 1) st4(io.psi_new[1] + fy + ((int64_t)(y + 1) * D2 + x), ny1);
          }
          // x faces x..x+4 of row y
          float ox[5], nx[5];
          faces5(oxv[k], oxe[k], own_edge, ox);
          const float pf[6] = {nb.y, pc.x, pc.y, pc.z, pc.w, nb.z};  // p at x-1 .. x+4
#pragma unroll
          for (int j = 0; j < 5; ++j)
            nx[j] =
                __builtin_fmaf((zcy - px.zf[j]) * px.k[j], pf[j + 1] - pf[j], px.a[j] * ox[j]);
          lap.x = lap.x + ((nx[1] + ox[1]) - (nx[0] + ox[0])) * ih2;
          lap.y = lap.y + ((nx[2] + ox[2]) - (nx[1] + ox[1])) * ih2;
          lap.z = lap.z + ((nx[3] + ox[3]) - (nx[2] + ox[2])) * ih2;
          lap.w = lap.w + ((nx[4] + ox[4]) - (nx[3] + ox[3])) * ih2;
          if (!lq) {
            float* nx_row = io.psi_new[2] + fx + (int64_t)y * F2 + x;
            st4(nx_row, make_float4(nx[0], nx[1], nx[2], nx[3]));
            if (last_col) nx_row[4] = nx[4];
          }
          if (io.ck_psi[1] && !lq) {
            st4(io.ck_psi[1] + fy + ((int64_t)y * D2 + x), oy0);
            if (y == D1 - 1) st4(io.ck_psi[1] + fy + ((int64_t)(y + 1) * D2 + x), oy1);
            float* cx = io.ck_psi[2] + fx + (int64_t)y * F2 + x;
            st4(cx, make_float4(ox[0], ox[1], ox[2], ox[3]));
            if (last_col) cx[4] = ox[4];
          }
        }
        // the Laplacian with both face divergences: the reverse step's imaging operand
        if (io.lap_out) st4(io.lap_out + base + (y * P2 + x), lap);
        Damp4<NW_PML> dm;
        if constexpr (Q) {
          dm.a = dm.b = dm.tap = f4(1.f);
        } else {
          dm = damp4_row<NW_PML>(ldc(g.damp[1], y), hx, dcol);
        }
        const int ro = y * P2 + x;
        const int tpv = io.taper_prev;
        float4 o = make_float4(
            rhs1<NW_PML>(pc.x, pp[k].x, mm[k].x, lap.x, dm.b.x, dm.tap.x, tpv),
            rhs1<NW_PML>(pc.y, pp[k].y, mm[k].y, lap.y, dm.b.y, dm.tap.y, tpv),
            rhs1<NW_PML>(pc.z, pp[k].z, mm[k].z, lap.z, dm.b.z, dm.tap.z, tpv),
            rhs1<NW_PML>(pc.w, pp[k].w, mm[k].w, lap.w, dm.b.w, dm.tap.w, tpv));
        if constexpr (!Q) {
          const float zzr = g.dt2 * zcy;
          o.x = o.x - (zzr * zcx.x) * pc.x;
          o.y = o.y - (zzr * zcx.y) * pc.y;
          o.z = o.z - (zzr * zcx.z) * pc.z;
          o.w = o.w - (zzr * zcx.w) * pc.w;
        }
        for (int t = 0; t < io.src_taps; ++t) {
          const int d = ldc(io.src_idx, s * io.src_taps + t) - ro;
          if (d >= 0 && d < 4) {
            const float sv = ldc(io.src_val, s * io.src_taps + t) * w_n;
            o.x += d == 0 ? sv : 0.f;
            o.y += d == 1 ? sv : 0.f;
            o.z += d == 2 ? sv : 0.f;
            o.w += d == 3 ? sv : 0.f;
          }
        }
        if constexpr (!Q) o = mul4(dm.a, o);
        st4(io.p_next + base + ro, o);
        if (io.ck_prev) {
          st4(io.ck_prev + base + ro, pp[k]);
          st4(io.ck_cur + base + ro, pc);
        }
      }
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) w[j] = w[j + 1];
    }
  }
}

template <int R, int TY>
__global__ void __launch_bounds__(256) k3_step_pml(Geo<float> g, StepIO<float> io, int shots,
                                                   int spw, int xcd_map) {
  using S = TileShape<TY, R>;
  __shared__ __align__(16) float lp[S::rows * kLW];
  const Tile tl =
      tile_of(div_up((int)g.D[2], 256), div_up((int)g.D[1], TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;  // uniform over the workgroup
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  if (y0 >= g.quiet[1][0] && y0 + TY <= g.quiet[1][1] && x0 >= g.quiet[2][0] &&
      x0 + 256 <= g.quiet[2][1])
    k3_step_pml_tile<R, TY, true>(g, io, shots, spw, tl, lp);
  else
    k3_step_pml_tile<R, TY, false>(g, io, shots, spw, tl, lp);
}

// One tile of k3_adjoint_pml (8 wavefronts, AdjShape rows each). Q: the tile lies in the PML's
// quiet box, where every face coefficient and the damping vanish: the adjoint face fields
// there enter lambda (here and in the neighbouring tiles) only multiplied by zero, so the tile
// neither reads nor writes them (they keep the finite values of the initial state) and steps
// with a = b = 1. The m a_eff tile is shot-independent and stays in registers; per shot every
// load (the lambda^{n+1} tile, the wave's lambda rows, kept Laplacians and the adjoint face
// fields its rows read) is issued before the first barrier, which orders LDS only.
template <int R, int TY, bool Q>
__device__ __forceinline__ void k3_adjoint_pml_tile(const Geo<float>& g, const AdjIO<float>& io,
                                                    int shots, int spw, const Tile& tl,
                                                    float* lq, float* lp) {
  using S = AdjShape<TY, R>;
  constexpr int MR = S::my_rows;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int F2 = (int)g.FP2;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  const int x = x0 + 4 * lane;
  const int r0 = wave * MR;
  const bool act = x < D2;
  const bool last_col = x + 4 == D2;
  const bool own_edge = lane == 63 || last_col;  // loads its face x+4 (see faces5)
  const float4 hx = act ? ld4(g.damp[2] + x) : f4(0.f);
  const Damp4<NW_PML> dcol = damp4_col<NW_PML>(hx);
  const float4 zcx = act ? ld4(g.zc[2] + x) : f4(0.f);
  const PmlX px = pml_x(g, x, act);
  const float ih1 = g.inv2h[1], ih2 = g.inv2h[2];
  const bool kept_lap = io.lap_n != nullptr;  // the forward step's Laplacian, faces included
  const bool imaging = (io.p_n != nullptr || kept_lap) && io.grad != nullptr;
  const bool p_tile = imaging && !kept_lap;
  // lanes whose four cells lie in the quiet box skip their face loads and stores as well
  const bool col_quiet = x >= g.quiet[2][0] && x + 4 <= g.quiet[2][1];
  float4 vm[S::per_thread];
  tile_load_w<S, R>(io.ma, y0, x0, D1, D2, P2, vm);
  double2 ga[MR], gb[MR];
#pragma unroll
  for (int k = 0; k < MR; ++k) {
    ga[k] = make_double2(0.0, 0.0);
    gb[k] = ga[k];
  }
  const int s_end = min(shots, (tl.grp + 1) * spw);
  for (int s = tl.grp * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const int64_t fy = (int64_t)s * g.fcells[1], fx = (int64_t)s * g.fcells[2];
    const float* L1 = io.lam1 + base;
    float4 vl[S::per_thread], vp[S::per_thread];
    tile_load_w<S, R>(L1, y0, x0, D1, D2, P2, vl);
    if (p_tile) tile_load_w<S, R>(io.p_n + base, y0, x0, D1, D2, P2, vp);
    // this wave's rows: lambda^{n+1}, lambda^{n+2}, the kept Laplacian and the adjoint faces
    float4 lc[MR], l2c[MR], lpn[MR], i0[MR], i1[MR], ix[MR];
    float ixe[MR];
#pragma unroll
    for (int k = 0; k < MR; ++k) {
      const int y = y0 + r0 + k;
      const bool ok = act && y < D1;
      lc[k] = ok ? ld4(L1 + (y * P2 + x)) : f4(0.f);
      l2c[k] = ok ? ld4nt(io.lam2 + base + (y * P2 + x)) : f4(0.f);
      lpn[k] = (ok && kept_lap && imaging) ? ld4(io.lap_n + base + (y * P2 + x)) : f4(0.f);
      const bool faces = !Q && ok && !(col_quiet && y >= g.quiet[1][0] && y < g.quiet[1][1]);
      i0[k] = faces ? ld4(io.psib_in[1] + fy + ((int64_t)y * D2 + x)) : f4(0.f);
      i1[k] = faces ? ld4(io.psib_in[1] + fy + ((int64_t)(y + 1) * D2 + x)) : f4(0.f);
      // x faces x..x+3, and x+4 where the next lane does not hold it
      const float* in_row = io.psib_in[2] + fx + (int64_t)y * F2 + x;
      ix[k] = faces ? ld4nt(in_row) : f4(0.f);  // read once: streaming
      ixe[k] = (faces && own_edge && x + 4 <= D2) ? in_row[4] : 0.f;
    }
    lds_barrier();  // the previous shot's reads of lq / lp are done
#pragma unroll
    for (int k = 0; k < S::per_thread; ++k) vl[k] = mul4(vm[k], vl[k]);
    tile_store_w<S>(lq, vl);
    if (p_tile) tile_store_w<S>(lp, vp);
    lds_barrier();
    float4 qw[2 * R + 1], pw[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) {
      qw[k] = lrow(lq, r0 + k, lane);
      pw[k] = p_tile ? lrow(lp, r0 + k, lane) : f4(0.f);
    }
#pragma unroll
    for (int k = 0; k < MR; ++k) {
      const int y = y0 + r0 + k;
      qw[2 * R] = lrow(lq, r0 + k + 2 * R, lane);
      pw[2 * R] = p_tile ? lrow(lp, r0 + k + 2 * R, lane) : f4(0.f);
      if (act && y < D1) {
        const int ro = y * P2 + x;
        const float4 qn = lnb(lq, r0 + k + R, lane);
        const float4 lapq = stencil4<R>(g, qw, qn.x, qn.y, qn.z, qn.w);
        Damp4<NW_PML> dm;
        if constexpr (Q) {
          dm.a = dm.b = dm.tap = f4(1.f);
        } else {
          dm = damp4_row<NW_PML>(ldc(g.damp[1], y), hx, dcol);
        }
        const float4 lcv = lc[k], l2v = l2c[k];
        const float zcy = Q ? 0.f : ldc(g.zc[1], y);
        const float zzr = g.dt2 * zcy;
        float4 rb, lam;
#define NW_PML_ADJ(c)                                                                        \
  {                                                                                          \
    rb.c = dm.a.c * lcv.c;                                                                   \
    const float pbar = -dm.b.c * (dm.a.c * l2v.c);                                           \
    lam.c = __builtin_fmaf(2.f - zzr * zcx.c, rb.c, pbar) + lapq.c;                          \
  }
        NW_PML_ADJ(x) NW_PML_ADJ(y) NW_PML_ADJ(z) NW_PML_ADJ(w)
#undef NW_PML_ADJ
        float4 lapp = f4(0.f);
        if (kept_lap && imaging) {
          lapp = lpn[k];
        } else if (p_tile) {
          const float4 pn = lnb(lp, r0 + k + R, lane);
          lapp = stencil4<R>(g, pw, pn.x, pn.y, pn.z, pn.w);
        }
        if constexpr (!Q) {
          const bool lqr = col_quiet && y >= g.quiet[1][0] && y < g.quiet[1][1];
          // y faces of rows y and y + 1
          {
            const float4 qc = qw[R], qm = qw[R - 1], qp = qw[R + 1];
            const float ay0 = ldc(g.alpha[1], y), ay1 = ldc(g.alpha[1], y + 1);
            const float ky0 = ldc(g.kappa[1], y), ky1 = ldc(g.kappa[1], y + 1);
            const float zy0 = ldc(g.zf[1], y), zy1 = ldc(g.zf[1], y + 1);
            const float4 a0 = i0[k], a1 = i1[k];
            float4 o0, o1;
#define NW_PML_ADJ_Y(c)                                                                      \
  {                                                                                          \
    const float dq0 = qc.c - qm.c, dq1 = qp.c - qc.c;                                        \
    const float t0 = a0.c - ih1 * dq0, t1 = a1.c - ih1 * dq1;                                \
    const float b0 = (zcx.c - zy0) * ky0, b1 = (zcx.c - zy1) * ky1;                          \
    lam.c = lam.c - (b1 * t1 - b0 * t0);                                                     \
    o0.c = ay0 * t0 - ih1 * dq0;                                                             \
    o1.c = ay1 * t1 - ih1 * dq1;                                                             \
  }
            NW_PML_ADJ_Y(x) NW_PML_ADJ_Y(y) NW_PML_ADJ_Y(z) NW_PML_ADJ_Y(w)
#undef NW_PML_ADJ_Y
            if (!lqr) {
              st4(io.psib_out[1] + fy + ((int64_t)y * D2 + x), o0);
              if (y == D1 - 1) st4(io.psib_out[1] + fy + ((int64_t)(y + 1) * D2 + x), o1);
            }
            if (p_tile && !lqr) {
              const float4 h0 = ld4(io.psi_half[1] + fy + ((int64_t)y * D2 + x));
              const float4 h1 = ld4(io.psi_half[1] + fy + ((int64_t)(y + 1) * D2 + x));
              const float4 w0 = ld4(io.psi_low[1] + fy + ((int64_t)y * D2 + x));
              const float4 w1 = ld4(io.psi_low[1] + fy + ((int64_t)(y + 1) * D2 + x));
              lapp.x = lapp.x + ((h1.x + w1.x) - (h0.x + w0.x)) * ih1;
              lapp.y = lapp.y + ((h1.y + w1.y) - (h0.y + w0.y)) * ih1;
              lapp.z = lapp.z + ((h1.z + w1.z) - (h0.z + w0.z)) * ih1;
              lapp.w = lapp.w + ((h1.w + w1.w) - (h0.w + w0.w)) * ih1;
            }
          }
          // x faces x..x+4 of row y. Face x+4 of a lane whose right neighbour skipped its
          // faces (quiet) reads as zero instead of its stored value: its coefficient b[4] is
          // zero there, so only the sign of a zero product can differ.
          {
            const float4 qc = qw[R];
            const float qf[6] = {qn.y, qc.x, qc.y, qc.z, qc.w, qn.z};  // q at x-1 .. x+4
            float inx[5], t[5], dq[5], b[5];
            faces5(ix[k], ixe[k], own_edge, inx);
#pragma unroll
            for (int j = 0; j < 5; ++j) {
              dq[j] = qf[j + 1] - qf[j];
              t[j] = inx[j] - ih2 * dq[j];
              b[j] = (zcy - px.zf[j]) * px.k[j];
            }
            lam.x = lam.x - (b[1] * t[1] - b[0] * t[0]);
            lam.y = lam.y - (b[2] * t[2] - b[1] * t[1]);
            lam.z = lam.z - (b[3] * t[3] - b[2] * t[2]);
            lam.w = lam.w - (b[4] * t[4] - b[3] * t[3]);
            float* out_row = io.psib_out[2] + fx + (int64_t)y * F2 + x;
            if (!lqr) {
              st4(out_row, make_float4(px.a[0] * t[0] - ih2 * dq[0], px.a[1] * t[1] - ih2 * dq[1],
                                       px.a[2] * t[2] - ih2 * dq[2],
                                       px.a[3] * t[3] - ih2 * dq[3]));
              if (last_col) out_row[4] = px.a[4] * t[4] - ih2 * dq[4];
            }
            if (p_tile) {
              // psi^{n+1/2} + psi^{n-1/2} at faces x..x+4 (quiet lanes pass zeros, the forward
              // face fields' values in the box, to the shuffle and add nothing themselves)
              const float* hr = io.psi_half[2] + fx + (int64_t)y * F2 + x;
              const float* wr = io.psi_low[2] + fx + (int64_t)y * F2 + x;
              const float4 h4 = lqr ? f4(0.f) : ld4nt(hr), w4 = lqr ? f4(0.f) : ld4nt(wr);
              const bool e = !lqr && own_edge && x + 4 <= D2;
              float hw[5];
              faces5(make_float4(h4.x + w4.x, h4.y + w4.y, h4.z + w4.z, h4.w + w4.w),
                     e ? hr[4] + wr[4] : 0.f, own_edge, hw);
              if (!lqr) {
                lapp.x = lapp.x + (hw[1] - hw[0]) * ih2;
                lapp.y = lapp.y + (hw[2] - hw[1]) * ih2;
                lapp.z = lapp.z + (hw[3] - hw[2]) * ih2;
                lapp.w = lapp.w + (hw[4] - hw[3]) * ih2;
              }
            }
          }
        }
        st4(io.lam0 + base + ro, lam);
        if (imaging) {
          ga[k].x = __builtin_fma((double)rb.x, (double)lapp.x, ga[k].x);
          ga[k].y = __builtin_fma((double)rb.y, (double)lapp.y, ga[k].y);
          gb[k].x = __builtin_fma((double)rb.z, (double)lapp.z, gb[k].x);
          gb[k].y = __builtin_fma((double)rb.w, (double)lapp.w, gb[k].y);
        }
      }
#pragma unroll
      for (int j = 0; j < 2 * R; ++j) {
        qw[j] = qw[j + 1];
        pw[j] = pw[j + 1];
      }
    }
  }
  if (imaging) {
    double* G = io.grad + (int64_t)tl.grp * g.cells;
#pragma unroll
    for (int k = 0; k < MR; ++k) {
      const int y = y0 + r0 + k;
      if (act && y < D1) {
        double2* gp = reinterpret_cast<double2*>(G + (y * P2 + x));
        double2 a = gp[0], b = gp[1];
        a.x += ga[k].x;
        a.y += ga[k].y;
        b.x += gb[k].x;
        b.y += gb[k].y;
        gp[0] = a;
        gp[1] = b;
      }
    }
  }
}

template <int R, int TY>
__global__ void __launch_bounds__(64 * AdjWaves<TY>::waves) k3_adjoint_pml(Geo<float> g,
                                                                          AdjIO<float> io,
                                                                          int shots, int spw,
                                                                          int xcd_map) {
  using S = AdjShape<TY, R>;
  __shared__ __align__(16) float lq[S::rows * kLW];  // q = (m a_eff) lambda^{n+1}
  __shared__ __align__(16) float lp[S::rows * kLW];  // p^n
  const Tile tl =
      tile_of(div_up((int)g.D[2], 256), div_up((int)g.D[1], TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;  // uniform over the workgroup
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  if (y0 >= g.quiet[1][0] && y0 + TY <= g.quiet[1][1] && x0 >= g.quiet[2][0] &&
      x0 + 256 <= g.quiet[2][1])
    k3_adjoint_pml_tile<R, TY, true>(g, io, shots, spw, tl, lq, lp);
  else
    k3_adjoint_pml_tile<R, TY, false>(g, io, shots, spw, tl, lq, lp);
}

// ------------------------------------------------------------------------------------------
// 3-D steps: 2.5-D tiles marching along the slowest axis
// ------------------------------------------------------------------------------------------
// A workgroup owns TY rows x 256 columns of one shot's grid and a chunk of ZC planes of the
// slow axis (axis 0). The planes of the stencil's input field sit in a ring of 2R+1 LDS planes
// (each with R halo rows and 4 margin columns on either side); plane z is computed from the
// ring while the loads of plane z+R+1 are in flight in registers, so each workgroup waits on
// HBM only in its prologue. Each input plane is read once per workgroup plus the chunk's 2R
// halo planes and the strip's 2R halo rows (both mostly from L2); the pointwise operands
// (p^{n-1}, m; lambda^{n+1}, lambda^{n+2}) are read per cell. Per cell the arithmetic is the
// 2-D tiles' with axis 0 in front: centre term, then the fused multiply-adds of axes 0, 1, 2;
// damping from the three profile values summed (sponge) or minimised (Cerjan) in axis order.
template <int TY, int R>
struct Shape3 {
  static constexpr int rows = TY + 2 * R;
  static constexpr int n4 = rows * (kLW / 4);
  static constexpr int per_thread = (n4 + 255) / 256;
  static constexpr int my_rows = TY / 4;
  static constexpr int ring = 2 * R + 1;
  static constexpr int plane = rows * kLW;  // floats of one LDS plane
};

struct Tile3 {
  int cx, ry, cz, s;
  bool valid;
};

// tiles numbered column chunk fastest, then row strip, z chunk, shot; XCD-contiguous ranges
// as in tile_of
__device__ __forceinline__ Tile3 tile3_of(int nX, int nY, int nZ, int shots, int xcd_map) {
  const int T = nX * nY * nZ * shots;
  const int per = (T + kXcds - 1) / kXcds;
  const int i = blockIdx.x;
  const int t = xcd_map ? (i % kXcds) * per + i / kXcds : i;
  Tile3 r;
  r.valid = t < T;
  r.cx = t % nX;
  r.ry = (t / nX) % nY;
  r.cz = (t / (nX * nY)) % nZ;
  r.s = t / (nX * nY * nZ);
  return r;
}

// damping of one cell from its three profile values
template <int ABS>
__device__ __forceinline__ void damp3(float hz, float hy, float hx, float& a, float& b,
                                      float& tap) {
  a = 1.f;
  b = 1.f;
  tap = 1.f;
  if (ABS == NW_SPONGE) {
    const float half = (hz + hy) + hx;
    a = 1.f / (1.f + half);
    b = 1.f - half;
  } else if (ABS == NW_CERJAN) {
    tap = hz < tap ? hz : tap;
    tap = hy < tap ? hy : tap;
    tap = hx < tap ? hx : tap;
    a = tap;
  }
}

template <int ABS>
__device__ __forceinline__ Damp4<ABS> damp3_4(float hz, float hy, const float4& hx) {
  Damp4<ABS> d;
  damp3<ABS>(hz, hy, hx.x, d.a.x, d.b.x, d.tap.x);
  damp3<ABS>(hz, hy, hx.y, d.a.y, d.b.y, d.tap.y);
  damp3<ABS>(hz, hy, hx.z, d.a.z, d.b.z, d.tap.z);
  damp3<ABS>(hz, hy, hx.w, d.a.w, d.b.w, d.tap.w);
  return d;
}

// one cell of the 3-D cross stencil
template <int R>
__device__ __forceinline__ float stencil3(const Geo<float>& g, float c, float zm1, float zp1,
                                          float zm2, float zp2, float ym1, float yp1, float ym2,
                                          float yp2, float xm1, float xp1, float xm2, float xp2) {
  float s = __builtin_fmaf(g.c1[0], zm1 + zp1, g.cc * c);
  if (R == 2) s = __builtin_fmaf(g.c2[0], zm2 + zp2, s);
  s = __builtin_fmaf(g.c1[1], ym1 + yp1, s);
  if (R == 2) s = __builtin_fmaf(g.c2[1], ym2 + yp2, s);
  s = __builtin_fmaf(g.c1[2], xm1 + xp1, s);
  if (R == 2) s = __builtin_fmaf(g.c2[2], xm2 + xp2, s);
  return s;
}

// stencil at the lane's four columns of LDS row r of the ring plane pz; the slow-axis
// neighbours from the ring planes pm1, pp1 (and pm2, pp2 for R = 2) at the same row
template <int R>
__device__ __forceinline__ float4 lap3(const Geo<float>& g, const float* pz, const float* pm1,
                                       const float* pp1, const float* pm2, const float* pp2,
                                       int r, int lane) {
  const float4 c = lrow(pz, r, lane);
  const float4 zm1 = lrow(pm1, r, lane), zp1 = lrow(pp1, r, lane);
  const float4 zm2 = R == 2 ? lrow(pm2, r, lane) : f4(0.f);
  const float4 zp2 = R == 2 ? lrow(pp2, r, lane) : f4(0.f);
  const float4 ym1 = lrow(pz, r - 1, lane), yp1 = lrow(pz, r + 1, lane);
  const float4 ym2 = R == 2 ? lrow(pz, r - 2, lane) : f4(0.f);
  const float4 yp2 = R == 2 ? lrow(pz, r + 2, lane) : f4(0.f);
  const float4 nb = lnb(pz, r, lane);  // x-2, x-1, x+4, x+5
  float4 o;
  o.x = stencil3<R>(g, c.x, zm1.x, zp1.x, zm2.x, zp2.x, ym1.x, yp1.x, ym2.x, yp2.x, nb.y, c.y,
                    nb.x, c.z);
  o.y = stencil3<R>(g, c.y, zm1.y, zp1.y, zm2.y, zp2.y, ym1.y, yp1.y, ym2.y, yp2.y, c.x, c.z,
                    nb.y, c.w);
  o.z = stencil3<R>(g, c.z, zm1.z, zp1.z, zm2.z, zp2.z, ym1.z, yp1.z, ym2.z, yp2.z, c.y, c.w,
                    c.x, nb.z);
  o.w = stencil3<R>(g, c.w, zm1.w, zp1.w, zm2.w, zp2.w, ym1.w, yp1.w, ym2.w, yp2.w, c.z, nb.z,
                    c.y, nb.w);
  return o;
}

// Loads of plane z of a field's tile (rows y0-R .. y0+TY+R-1, columns x0-4 .. x0+259; zero
// outside the grid), multiplied pointwise by the same cells of `mul` when it is given.
template <int TY, int R>
__device__ __forceinline__ void plane_fetch(const float* f, const float* mul, int z, int y0,
                                            int x0, int D0, int D1, int D2, int P2,
                                            int64_t plane, float4* v) {
  using S = Shape3<TY, R>;
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int row = i / (kLW / 4), c4 = i % (kLW / 4);
    const int gy = y0 - R + row, gx = x0 - 4 + 4 * c4;
    const bool ok = i < S::n4 && z >= 0 && z < D0 && gy >= 0 && gy < D1 && gx >= 0 && gx < D2;
    const int64_t off = (int64_t)z * plane + (gy * P2 + gx);
    v[k] = ok ? ld4(f + off) : f4(0.f);
    if (mul != nullptr && ok) v[k] = mul4(ld4(mul + off), v[k]);
  }
}

template <int TY, int R>
__device__ __forceinline__ void plane_put(float* lds, const float4* v) {
  using S = Shape3<TY, R>;
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    if (i < S::n4) *reinterpret_cast<float4*>(lds + 4 * i) = v[k];
  }
}

__device__ __forceinline__ int ring_slot(int z, int n) { return ((z % n) + n) % n; }

// The ring march shared by both kernels: fills the ring with planes z0-R .. z0+R-1, then for
// each plane z of the chunk puts plane z+R into the ring, starts the loads of plane z+R+1
// (calling fetched(z+R+1) right after them) and calls cell(z, ring planes) once the ring holds
// z-R .. z+R.
template <int TY, int R, typename F, typename G>
__device__ __forceinline__ void ring_march(float* ring, const float* f, const float* mul,
                                           int z0, int z1, int y0, int x0, int D0, int D1,
                                           int D2, int P2, int64_t plane, F cell, G fetched) {
  using S = Shape3<TY, R>;
  float4 v[S::per_thread];
  for (int z = z0 - R; z < z0 + R; ++z) {
    plane_fetch<TY, R>(f, mul, z, y0, x0, D0, D1, D2, P2, plane, v);
    plane_put<TY, R>(ring + ring_slot(z, S::ring) * S::plane, v);
  }
  plane_fetch<TY, R>(f, mul, z0 + R, y0, x0, D0, D1, D2, P2, plane, v);
  for (int z = z0; z < z1; ++z) {
    __syncthreads();  // plane z-1 is computed: the slot of plane z-R-1 is free
    plane_put<TY, R>(ring + ring_slot(z + R, S::ring) * S::plane, v);
    __syncthreads();
    if (z + 1 < z1) {
      plane_fetch<TY, R>(f, mul, z + R + 1, y0, x0, D0, D1, D2, P2, plane, v);
      fetched(z + R + 1);
    }
    const float* pz = ring + ring_slot(z, S::ring) * S::plane;
    const float* pm1 = ring + ring_slot(z - 1, S::ring) * S::plane;
    const float* pp1 = ring + ring_slot(z + 1, S::ring) * S::plane;
    const float* pm2 = R == 2 ? ring + ring_slot(z - 2, S::ring) * S::plane : pz;
    const float* pp2 = R == 2 ? ring + ring_slot(z + 2, S::ring) * S::plane : pz;
    cell(z, pz, pm1, pp1, pm2, pp2);
  }
}

// Receiver gather of p^n by the workgroups that follow a step launch's tiles: workgroup b of
// them takes shot b / per_shot and the receiver taps 256 (b % per_shot) .. +255 of that shot,
// with k_sample's arithmetic (tap products summed by the same shuffle butterfly), so a step
// samples its own input p^n without a launch of its own.
__device__ __forceinline__ void sample_block(const Geo<float>& g, const SampleIO<float>& io,
                                             int b, int per_shot) {
  const int s = b / per_shot;
  const int t = (b % per_shot) * 256 + threadIdx.x;
  const int taps = io.taps;
  const int r = t / taps, k = t % taps;
  const int count = io.rec_count[s];
  float v = 0.f;
  if (r < count) {
    const int64_t row = (int64_t)io.rec_first[s] + r;
    v = io.rec_w[row * taps + k] * io.p[(int64_t)s * g.cells + io.rec_idx[row * taps + k]];
  }
  for (int off = 1; off < taps; off <<= 1) v += __shfl_xor(v, off, 64);
  if (r < count && k == 0) io.traces[((int64_t)io.trace_first[s] + r) * io.n_t + io.j] = v;
}

template <int R, int ABS, int TY, int ZC>
__global__ void __launch_bounds__(256) k3d_step(Geo<float> g, StepIO<float> io, int shots,
                                                int xcd_map, SampleIO<float> smp,
                                                int tile_blocks, int sample_per_shot) {
  using S = Shape3<TY, R>;
  __shared__ __align__(16) float ring[S::ring * S::plane];
  if ((int)blockIdx.x >= tile_blocks) {  // uniform over the workgroup
    sample_block(g, smp, (int)blockIdx.x - tile_blocks, sample_per_shot);
    return;
  }
  const int D0 = (int)g.D[0], D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int64_t plane = (int64_t)D1 * P2;
  const Tile3 tl = tile3_of(div_up(D2, 256), div_up(D1, TY), div_up(D0, ZC), shots, xcd_map);
  if (!tl.valid) return;  // uniform over the workgroup
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int x0 = tl.cx * 256, y0 = tl.ry * TY, z0 = tl.cz * ZC, s = tl.s;
  const int x = x0 + 4 * lane;
  const int r0 = wave * S::my_rows;
  const bool act = x < D2;
  const int64_t base = (int64_t)s * g.cells;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const float w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.f;
  ring_march<TY, R>(ring, io.p + base, nullptr, z0, min(D0, z0 + ZC), y0, x0, D0, D1, D2, P2,
                    plane, [&](int z, const float* pz, const float* pm1, const float* pp1,
                               const float* pm2, const float* pp2) {
    const float hz = ABS != NW_NONE ? g.damp[0][z] : 0.f;
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      if (!act || y >= D1) continue;
      const int rr = r0 + k + R;
      const float4 lap = lap3<R>(g, pz, pm1, pp1, pm2, pp2, rr, lane);
      const int64_t ro = (int64_t)z * plane + (y * P2 + x);
      const float4 pc = lrow(pz, rr, lane);
      const float4 pp = ld4(io.p_prev + base + ro);
      const float4 mm = ld4(io.m + ro);
      const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
      const Damp4<ABS> dm = damp3_4<ABS>(hz, hy, hx);
      const int tpv = io.taper_prev;
      float4 o = make_float4(rhs1<ABS>(pc.x, pp.x, mm.x, lap.x, dm.b.x, dm.tap.x, tpv),
                             rhs1<ABS>(pc.y, pp.y, mm.y, lap.y, dm.b.y, dm.tap.y, tpv),
                             rhs1<ABS>(pc.z, pp.z, mm.z, lap.z, dm.b.z, dm.tap.z, tpv),
                             rhs1<ABS>(pc.w, pp.w, mm.w, lap.w, dm.b.w, dm.tap.w, tpv));
      for (int t = 0; t < io.src_taps; ++t) {
        const int64_t d = ldc(io.src_idx, s * io.src_taps + t) - ro;
        if (d >= 0 && d < 4) {
          const float sv = ldc(io.src_val, s * io.src_taps + t) * w_n;
          o.x += d == 0 ? sv : 0.f;
          o.y += d == 1 ? sv : 0.f;
          o.z += d == 2 ? sv : 0.f;
          o.w += d == 3 ? sv : 0.f;
        }
      }
      if (ABS != NW_NONE) o = mul4(dm.a, o);
      st4p(io.p_next + base + ro, o);
      if (io.lap_out) st4p(io.lap_out + base + ro, lap);
      if (io.ck_prev) {
        st4p(io.ck_prev + base + ro, pp);
        st4p(io.ck_cur + base + ro, pc);
      }
    }
  }, [](int) {});
}

// The reverse step: lambda^n from the stencil of q = (m a) lambda^{n+1}, which the ring holds,
// and the pointwise lambda^{n+1}, lambda^{n+2}. The imaging sum, when there is one, reads the
// Laplacian of p^n the forward sweep kept (io.lap_n) and adds a lambda^{n+1} * lap p^n to the
// shot's float64 gradient slice; the initial-value operator's sweep has none.
template <int R, int ABS, int TY, int ZC>
__global__ void __launch_bounds__(256) k3d_adjoint(Geo<float> g, AdjIO<float> io, int shots,
                                                   int xcd_map) {
  using S = Shape3<TY, R>;
  __shared__ __align__(16) float ring[S::ring * S::plane];
  const int D0 = (int)g.D[0], D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int64_t plane = (int64_t)D1 * P2;
  const Tile3 tl = tile3_of(div_up(D2, 256), div_up(D1, TY), div_up(D0, ZC), shots, xcd_map);
  if (!tl.valid) return;
  const int lane = threadIdx.x & 63, wave = wave_index();
  const int x0 = tl.cx * 256, y0 = tl.ry * TY, z0 = tl.cz * ZC, s = tl.s;
  const int x = x0 + 4 * lane;
  const int r0 = wave * S::my_rows;
  const bool act = x < D2;
  const int64_t base = (int64_t)s * g.cells;
  const float* L1 = io.lam1 + base;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const int tpn = io.taper_next;
  const int z1 = min(D0, z0 + ZC);
  const bool imaging = io.lap_n != nullptr && io.grad != nullptr;
  // lambda^{n+1} at the lane's own cells of planes z .. z+R+1 (lq[0] = plane z), loaded beside
  // the ring's loads of the same plane so that both are served by one trip to memory
  float4 lq[R + 2][S::my_rows];
  auto own = [&](int zz, float4* out) {
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      out[k] = (act && y < D1 && zz < D0) ? ld4(L1 + ((int64_t)zz * plane + (y * P2 + x)))
                                          : f4(0.f);
    }
  };
#pragma unroll
  for (int j = 0; j <= R; ++j) own(z0 + j, lq[j]);
  ring_march<TY, R>(ring, L1, io.ma, z0, z1, y0, x0, D0, D1, D2, P2, plane,
                    [&](int z, const float* pz, const float* pm1, const float* pp1,
                        const float* pm2, const float* pp2) {
    const float hz = ABS != NW_NONE ? g.damp[0][z] : 0.f;
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      if (!act || y >= D1) continue;
      const int rr = r0 + k + R;
      const float4 lapq = lap3<R>(g, pz, pm1, pp1, pm2, pp2, rr, lane);
      const int64_t ro = (int64_t)z * plane + (y * P2 + x);
      const float4 lc = lq[0][k], l2c = ld4(io.lam2 + base + ro);
      const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
      const Damp4<ABS> dm = damp3_4<ABS>(hz, hy, hx);
      float4 rb;
      const float4 o = make_float4(
          adj1<ABS>(lc.x, l2c.x, lapq.x, dm.a.x, dm.b.x, dm.tap.x, tpn, rb.x),
          adj1<ABS>(lc.y, l2c.y, lapq.y, dm.a.y, dm.b.y, dm.tap.y, tpn, rb.y),
          adj1<ABS>(lc.z, l2c.z, lapq.z, dm.a.z, dm.b.z, dm.tap.z, tpn, rb.z),
          adj1<ABS>(lc.w, l2c.w, lapq.w, dm.a.w, dm.b.w, dm.tap.w, tpn, rb.w));
      st4p(io.lam0 + base + ro, o);
      if (imaging) {
        const float4 lapp = ld4(io.lap_n + base + ro);
        double2* gp = reinterpret_cast<double2*>(io.grad + base + ro);
        double2 ga = gp[0], gb = gp[1];
        ga.x = __builtin_fma((double)rb.x, (double)lapp.x, ga.x);
        ga.y = __builtin_fma((double)rb.y, (double)lapp.y, ga.y);
        gb.x = __builtin_fma((double)rb.z, (double)lapp.z, gb.x);
        gb.y = __builtin_fma((double)rb.w, (double)lapp.w, gb.y);
        gp[0] = ga;
        gp[1] = gb;
      }
    }
#pragma unroll
    for (int j = 0; j <= R; ++j)
#pragma unroll
      for (int k = 0; k < S::my_rows; ++k) lq[j][k] = lq[j + 1][k];
  }, [&](int zz) { own(zz, lq[R + 1]); });
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
bool fast2d_supported(const Geo<float>& g, int absorbing, int dtype) {
  (void)absorbing;
  return g.nd == 2 && dtype == NW_F32 && g.D[2] % 4 == 0 && g.P2 % 4 == 0;
}

// 1-D grid of kXcds * ceil(tiles / kXcds) workgroups (tile_of maps them, a few return)
static unsigned grid_tiles(const Geo<float>& g, int ty, int shots, int spw) {
  const int T = div_up((int)g.D[2], 256) * div_up((int)g.D[1], 4 * ty) * div_up(shots, spw);
  return (unsigned)(kXcds * div_up(T, kXcds));
}

static unsigned grid_tiles3(const Geo<float>& g, int ty, int shots, int spw) {
  const int T = div_up((int)g.D[2], 256) * div_up((int)g.D[1], ty) * div_up(shots, spw);
  return (unsigned)(kXcds * div_up(T, kXcds));
}

template <int TY>
constexpr int lds_rows() {
  return TY < 8 ? 8 : (TY > 32 ? 32 : TY);
}

// PML plans always step with the LDS-tiled face kernels, 8 or 16 rows (the 32-row variants
// need more than 256 VGPRs)
template <int TY>
constexpr int pml_rows() {
  return TY <= 8 ? 8 : 16;
}

template <int R, int ABS, int TY>
static void launch_step_t(const Geo<float>& g, const StepIO<float>& io, int shots, int spw,
                          int xm, hipStream_t st) {
  // xm bit 2 selects the LDS-tiled kernel, bit 1 the two-rows-per-iteration march
  if constexpr (ABS == NW_PML) {
    hipLaunchKernelGGL((k3_step_pml<R, pml_rows<TY>()>),
                       dim3(grid_tiles3(g, pml_rows<TY>(), shots, spw)), dim3(256), 0, st, g, io,
                       shots, spw, xm & 1);
  } else if ((xm & 4) && TY >= 8 && TY <= 32) {
    hipLaunchKernelGGL((k3_step<R, ABS, lds_rows<TY>()>), dim3(grid_tiles3(g, TY, shots, spw)),
                       dim3(256), 0, st, g, io, shots, spw, xm & 1);
  } else if (xm & 2) {
    hipLaunchKernelGGL((k2_step<R, ABS, TY, 2>), dim3(grid_tiles(g, TY, shots, spw)), dim3(256),
                       0, st, g, io, shots, spw, xm & 1);
  } else {
    hipLaunchKernelGGL((k2_step<R, ABS, TY, 1>), dim3(grid_tiles(g, TY, shots, spw)), dim3(256),
                       0, st, g, io, shots, spw, xm & 1);
  }
}

template <int R, int ABS, int TY>
static void launch_adj_t(const Geo<float>& g, const AdjIO<float>& io, int shots, int spw, int xm,
                         hipStream_t st) {
  if constexpr (ABS == NW_PML) {
    hipLaunchKernelGGL((k3_adjoint_pml<R, pml_rows<TY>()>),
                       dim3(grid_tiles3(g, pml_rows<TY>(), shots, spw)),
                       dim3(64 * AdjWaves<pml_rows<TY>()>::waves), 0, st, g, io,
                       shots, spw, xm & 1);
  } else if ((xm & 4) && TY >= 8 && TY <= 32) {
    const int tiles = block_tiles(adj_cols((int)g.D[2]), div_up((int)g.D[1], lds_rows<TY>()),
                                  div_up(shots, spw));
    hipLaunchKernelGGL((k3_adjoint<R, ABS, lds_rows<TY>()>),
                       dim3((unsigned)(kXcds * div_up(tiles, kXcds))),
                       dim3(64 * AdjWaves<lds_rows<TY>()>::waves), 0, st, g, io, shots, spw,
                       xm & 1);
  } else {
    hipLaunchKernelGGL((k2_adjoint<R, ABS, TY>), dim3(grid_tiles(g, TY, shots, spw)), dim3(256),
                       0, st, g, io, shots, spw, xm & 1);
  }
}

template <int R, int ABS>
static void step_ty(const Geo<float>& g, const StepIO<float>& io, int shots, int ty, int spw,
                    int xm, hipStream_t st) {
  switch (ty) {
    case 4: launch_step_t<R, ABS, 4>(g, io, shots, spw, xm, st); break;
    case 8: launch_step_t<R, ABS, 8>(g, io, shots, spw, xm, st); break;
    case 32: launch_step_t<R, ABS, 32>(g, io, shots, spw, xm, st); break;
    case 64: launch_step_t<R, ABS, 64>(g, io, shots, spw, xm, st); break;
    default: launch_step_t<R, ABS, 16>(g, io, shots, spw, xm, st); break;
  }
}

template <int R, int ABS>
static void adj_ty(const Geo<float>& g, const AdjIO<float>& io, int shots, int ty, int spw,
                   int xm, hipStream_t st) {
  switch (ty) {
    case 4: launch_adj_t<R, ABS, 4>(g, io, shots, spw, xm, st); break;
    case 8: launch_adj_t<R, ABS, 8>(g, io, shots, spw, xm, st); break;
    case 32: launch_adj_t<R, ABS, 32>(g, io, shots, spw, xm, st); break;
    case 64: launch_adj_t<R, ABS, 64>(g, io, shots, spw, xm, st); break;
    default: launch_adj_t<R, ABS, 16>(g, io, shots, spw, xm, st); break;
  }
}

void fast2d_step(const Geo<float>& g, int order, int absorbing, const StepIO<float>& io,
                 int shots, int ty, int spw, int xm, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int key = absorbing * 2 + (order == 4 ? 1 : 0);
  switch (key) {
    case NW_NONE * 2 + 0: step_ty<1, NW_NONE>(g, io, shots, ty, spw, xm, st); break;
    case NW_NONE * 2 + 1: step_ty<2, NW_NONE>(g, io, shots, ty, spw, xm, st); break;
    case NW_SPONGE * 2 + 0: step_ty<1, NW_SPONGE>(g, io, shots, ty, spw, xm, st); break;
    case NW_SPONGE * 2 + 1: step_ty<2, NW_SPONGE>(g, io, shots, ty, spw, xm, st); break;
    case NW_CERJAN * 2 + 0: step_ty<1, NW_CERJAN>(g, io, shots, ty, spw, xm, st); break;
    case NW_CERJAN * 2 + 1: step_ty<2, NW_CERJAN>(g, io, shots, ty, spw, xm, st); break;
    case NW_PML * 2 + 0: step_ty<1, NW_PML>(g, io, shots, ty, spw, xm, st); break;
    default: step_ty<2, NW_PML>(g, io, shots, ty, spw, xm, st); break;
  }
}

void fast2d_adjoint(const Geo<float>& g, int order, int absorbing, const AdjIO<float>& io,
                    int shots, int ty, int spw, int xm, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int key = absorbing * 2 + (order == 4 ? 1 : 0);
  switch (key) {
    case NW_NONE * 2 + 0: adj_ty<1, NW_NONE>(g, io, shots, ty, spw, xm, st); break;
    case NW_NONE * 2 + 1: adj_ty<2, NW_NONE>(g, io, shots, ty, spw, xm, st); break;
    case NW_SPONGE * 2 + 0: adj_ty<1, NW_SPONGE>(g, io, shots, ty, spw, xm, st); break;
    case NW_SPONGE * 2 + 1: adj_ty<2, NW_SPONGE>(g, io, shots, ty, spw, xm, st); break;
    case NW_CERJAN * 2 + 0: adj_ty<1, NW_CERJAN>(g, io, shots, ty, spw, xm, st); break;
    case NW_CERJAN * 2 + 1: adj_ty<2, NW_CERJAN>(g, io, shots, ty, spw, xm, st); break;
    case NW_PML * 2 + 0: adj_ty<1, NW_PML>(g, io, shots, ty, spw, xm, st); break;
    default: adj_ty<2, NW_PML>(g, io, shots, ty, spw, xm, st); break;
  }
}

bool fast2d_block_supported(int ty, int depth) {
  return (ty == 8 || ty == 16 || ty == 32) && depth >= 2 && depth <= kMaxDepth;
}

bool fast2d_adj_block_supported(int ty, int depth) {
  return (ty == 8 || ty == 16) && depth >= 2 && depth <= kMaxDepth;
}

// Calls fn(R, ABS, TY, T) with the compile-time shape of a blocked forward launch (integral
// constants), for the instances the library builds.
template <class Fn>
static void with_block_shape(int order, int absorbing, int ty, int depth, Fn&& fn) {
  auto by_depth = [&](auto r, auto abs, auto rows) {
    switch (depth) {
      case 2: fn(r, abs, rows, std::integral_constant<int, 2>{}); break;
      case 3: fn(r, abs, rows, std::integral_constant<int, 3>{}); break;
      default: fn(r, abs, rows, std::integral_constant<int, 4>{}); break;
    }
  };
  auto by_rows = [&](auto r, auto abs) {
    switch (ty) {
      case 8: by_depth(r, abs, std::integral_constant<int, 8>{}); break;
      case 32: by_depth(r, abs, std::integral_constant<int, 32>{}); break;
      default: by_depth(r, abs, std::integral_constant<int, 16>{}); break;
    }
  };
  using R1 = std::integral_constant<int, 1>;
  using R2 = std::integral_constant<int, 2>;
  switch (absorbing * 2 + (order == 4 ? 1 : 0)) {
    case NW_NONE * 2 + 0: by_rows(R1{}, std::integral_constant<int, NW_NONE>{}); break;
    case NW_NONE * 2 + 1: by_rows(R2{}, std::integral_constant<int, NW_NONE>{}); break;
    case NW_SPONGE * 2 + 0: by_rows(R1{}, std::integral_constant<int, NW_SPONGE>{}); break;
    case NW_SPONGE * 2 + 1: by_rows(R2{}, std::integral_constant<int, NW_SPONGE>{}); break;
    case NW_CERJAN * 2 + 0: by_rows(R1{}, std::integral_constant<int, NW_CERJAN>{}); break;
    case NW_CERJAN * 2 + 1: by_rows(R2{}, std::integral_constant<int, NW_CERJAN>{}); break;
    default: break;  // PML plans are never blocked (their face fields are not carried)
  }
}

void fast2d_steps(const Geo<float>& g, int order, int absorbing, const BlockIO& io, int shots,
                  int ty, int depth, int spw, int xm, const BlockQueue* queue, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  with_block_shape(order, absorbing, ty, depth, [&](auto r, auto abs, auto rows, auto t) {
    constexpr int R = decltype(r)::value, ABS = decltype(abs)::value;
    constexpr int TY = decltype(rows)::value, T = decltype(t)::value;
    using S = BlockShape<TY, R, T>;
    if (queue) {
      BlockQueues q;
      for (int j = 0; j <= kXcds; ++j) q.off[j] = queue->off[j];
      hipLaunchKernelGGL((k4_queue<R, ABS, TY, T>), dim3((unsigned)queue->grid),
                         dim3(S::threads), 0, st, g, io, shots, queue->items, q, queue->ctr);
    } else {
      const int tiles =
          block_tiles(block_cols<S>((int)g.D[2]), div_up((int)g.D[1], TY), div_up(shots, spw));
      hipLaunchKernelGGL((k4_steps<R, ABS, TY, T>),
                         dim3((unsigned)(kXcds * div_up(tiles, kXcds))), dim3(S::threads), 0,
                         st, g, io, shots, spw, xm & 1);
    }
  });
}

int fast2d_steps_queue_items(const Geo<float>& g, int order, int absorbing, int ty, int depth,
                             int shots, int spw, int split_tail, std::vector<int4>& items,
                             int* off, int* per_cu) {
  int rc = NW_ERR_ARG;
  with_block_shape(order, absorbing, ty, depth, [&](auto r, auto abs, auto rows, auto t) {
    constexpr int R = decltype(r)::value, ABS = decltype(abs)::value;
    constexpr int TY = decltype(rows)::value, T = decltype(t)::value;
    using S = BlockShape<TY, R, T>;
    const BlockCols cols = block_cols<S>((int)g.D[2]);
    const int nY = div_up((int)g.D[1], TY), nG = div_up(shots, spw);
    const int per_grp = block_tiles(cols, nY, 1), tiles = per_grp * nG;
    const int per = div_up(tiles, kXcds);
    items.clear();
    for (int j = 0; j < kXcds; ++j) {
      off[j] = (int)items.size();
      const int t0 = j * per, t1 = std::min(tiles, t0 + per), listed = t1 - t0;
      int k = 0;
      for (int tt = t0; tt < t1; ++tt, ++k) {
        // tile tt in tile_of_cols' order (without the XCD map, which these ranges replace)
        const int grp = tt / per_grp, u = tt - grp * per_grp;
        int cx, ry;
        if (!cols.split) {
          cx = u % cols.full;
          ry = u / cols.full;
        } else {
          const int pair = u / (2 * cols.full + 1), v = u - pair * (2 * cols.full + 1);
          cx = v < 2 * cols.full ? v % cols.full : cols.full;
          ry = v < 2 * cols.full ? 2 * pair + v / cols.full : 2 * pair;
        }
        if (ry >= nY) continue;  // the missing second strip of an odd last pair
        const int s0 = grp * spw, s1 = std::min(shots, s0 + spw);
        if (k >= listed - split_tail) {
          for (int s = s0; s < s1; ++s) items.push_back(make_int4(ry, cx, s, s + 1));
        } else {
          items.push_back(make_int4(ry, cx, s0, s1));
        }
      }
    }
    off[kXcds] = (int)items.size();
    int n = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, reinterpret_cast<const void*>(&k4_queue<R, ABS, TY, T>), S::threads, 0);
    *per_cu = e == hipSuccess ? std::max(n, 1) : 1;
    rc = e == hipSuccess ? NW_OK : NW_ERR_HIP;
  });
  return rc;
}

template <int R, int ABS, int TY, int T>
static void launch_adj_steps_t(const Geo<float>& g, const AdjBlockIO& io, int shots, int spw,
                               int xm, hipStream_t st) {
  using S = BlockShape<TY, R, T, AdjBlockWaves<TY>::waves>;
  const int tiles =
      div_up((int)g.D[2], S::out_w) * div_up((int)g.D[1], TY) * div_up(shots, spw);
  hipLaunchKernelGGL((k5_adjoint<R, ABS, TY, T>),
                     dim3((unsigned)(kXcds * div_up(tiles, kXcds))), dim3(S::threads), 0, st, g,
                     io, shots, spw, xm & 1);
}

template <int R, int ABS, int TY>
static void adj_steps_depth(const Geo<float>& g, const AdjBlockIO& io, int shots, int depth,
                            int spw, int xm, hipStream_t st) {
  switch (depth) {
    case 2: launch_adj_steps_t<R, ABS, TY, 2>(g, io, shots, spw, xm, st); break;
    case 3: launch_adj_steps_t<R, ABS, TY, 3>(g, io, shots, spw, xm, st); break;
    default: launch_adj_steps_t<R, ABS, TY, 4>(g, io, shots, spw, xm, st); break;
  }
}

template <int R, int ABS>
static void adj_steps_ty(const Geo<float>& g, const AdjBlockIO& io, int shots, int ty,
                         int depth, int spw, int xm, hipStream_t st) {
  if (ty == 8)
    adj_steps_depth<R, ABS, 8>(g, io, shots, depth, spw, xm, st);
  else
    adj_steps_depth<R, ABS, 16>(g, io, shots, depth, spw, xm, st);
}

void fast2d_adjoint_steps(const Geo<float>& g, int order, int absorbing, const AdjBlockIO& io,
                          int shots, int ty, int depth, int spw, int xm, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int key = absorbing * 2 + (order == 4 ? 1 : 0);
  switch (key) {
    case NW_NONE * 2 + 0: adj_steps_ty<1, NW_NONE>(g, io, shots, ty, depth, spw, xm, st); break;
    case NW_NONE * 2 + 1: adj_steps_ty<2, NW_NONE>(g, io, shots, ty, depth, spw, xm, st); break;
    case NW_SPONGE * 2 + 0: adj_steps_ty<1, NW_SPONGE>(g, io, shots, ty, depth, spw, xm, st); break;
    case NW_SPONGE * 2 + 1: adj_steps_ty<2, NW_SPONGE>(g, io, shots, ty, depth, spw, xm, st); break;
    case NW_CERJAN * 2 + 0: adj_steps_ty<1, NW_CERJAN>(g, io, shots, ty, depth, spw, xm, st); break;
    case NW_CERJAN * 2 + 1: adj_steps_ty<2, NW_CERJAN>(g, io, shots, ty, depth, spw, xm, st); break;
    default: break;  // PML plans are never blocked
  }
}

void fast2d_sample_levels(const Geo<float>& g, const SampleIO<float>& io, const float* tapbuf,
                          int tap_stride, int rec_max, int shots, int levels,
                          const LevelRecords& lj, void* stream) {
  if (levels < 2 || rec_max <= 0) return;
  const dim3 grid((unsigned)div_up(rec_max * io.taps, 256), (unsigned)shots,
                  (unsigned)(levels - 1));
  hipLaunchKernelGGL(k_sample_levels, grid, dim3(256), 0, (hipStream_t)stream, g, io, tapbuf,
                     tap_stride, shots, lj);
}


// ------------------------------------------------------------------------------------------
// 3-D launchers
// ------------------------------------------------------------------------------------------
bool fast3d_supported(const Geo<float>& g, int absorbing, int dtype) {
  return g.nd == 3 && dtype == NW_F32 && absorbing != NW_PML && g.D[2] % 4 == 0 &&
         g.P2 % 4 == 0;
}

// 4-row tiles unless 8-row tiles already give several workgroups per CU; 8-plane chunks
template <int TY>
static unsigned grid3(const Geo<float>& g, int shots) {
  const int T = div_up((int)g.D[2], 256) * div_up((int)g.D[1], TY) * div_up((int)g.D[0], 8) * shots;
  return (unsigned)(kXcds * div_up(T, kXcds));
}

static bool rows8(const Geo<float>& g, int shots) { return grid3<8>(g, shots) >= 2048; }

template <int R, int ABS>
static void step3_t(const Geo<float>& g, const StepIO<float>& io, int shots, int xm,
                    const SampleIO<float>& smp, int per_shot, hipStream_t st) {
  const int extra = per_shot * shots;  // receiver-gather workgroups after the tiles
  if (rows8(g, shots)) {
    const int tiles = (int)grid3<8>(g, shots);
    hipLaunchKernelGGL((k3d_step<R, ABS, 8, 8>), dim3(tiles + extra), dim3(256), 0, st, g, io,
                       shots, xm & 1, smp, tiles, per_shot);
  } else {
    const int tiles = (int)grid3<4>(g, shots);
    hipLaunchKernelGGL((k3d_step<R, ABS, 4, 8>), dim3(tiles + extra), dim3(256), 0, st, g, io,
                       shots, xm & 1, smp, tiles, per_shot);
  }
}

template <int R, int ABS>
static void adj3_t(const Geo<float>& g, const AdjIO<float>& io, int shots, int xm,
                   hipStream_t st) {
  if (rows8(g, shots))
    hipLaunchKernelGGL((k3d_adjoint<R, ABS, 8, 8>), dim3(grid3<8>(g, shots)), dim3(256), 0, st, g,
                       io, shots, xm & 1);
  else
    hipLaunchKernelGGL((k3d_adjoint<R, ABS, 4, 8>), dim3(grid3<4>(g, shots)), dim3(256), 0, st, g,
                       io, shots, xm & 1);
}

void fast3d_step(const Geo<float>& g, int order, int absorbing, const StepIO<float>& io, int shots,
                 int xm, const SampleIO<float>* sample, int rec_max, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int key = absorbing * 2 + (order == 4 ? 1 : 0);
  SampleIO<float> smp{};
  int per_shot = 0;
  if (sample && rec_max > 0) {
    smp = *sample;
    per_shot = div_up(rec_max * sample->taps, 256);
  }
  switch (key) {
    case NW_NONE * 2 + 0: step3_t<1, NW_NONE>(g, io, shots, xm, smp, per_shot, st); break;
    case NW_NONE * 2 + 1: step3_t<2, NW_NONE>(g, io, shots, xm, smp, per_shot, st); break;
    case NW_SPONGE * 2 + 0: step3_t<1, NW_SPONGE>(g, io, shots, xm, smp, per_shot, st); break;
    case NW_SPONGE * 2 + 1: step3_t<2, NW_SPONGE>(g, io, shots, xm, smp, per_shot, st); break;
    case NW_CERJAN * 2 + 0: step3_t<1, NW_CERJAN>(g, io, shots, xm, smp, per_shot, st); break;
    case NW_CERJAN * 2 + 1: step3_t<2, NW_CERJAN>(g, io, shots, xm, smp, per_shot, st); break;
    default: break;  // no 3-D PML
  }
}

void fast3d_adjoint(const Geo<float>& g, int order, int absorbing, const AdjIO<float>& io,
                    int shots, int xm, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int key = absorbing * 2 + (order == 4 ? 1 : 0);
  switch (key) {
    case NW_NONE * 2 + 0: adj3_t<1, NW_NONE>(g, io, shots, xm, st); break;
    case NW_NONE * 2 + 1: adj3_t<2, NW_NONE>(g, io, shots, xm, st); break;
    case NW_SPONGE * 2 + 0: adj3_t<1, NW_SPONGE>(g, io, shots, xm, st); break;
    case NW_SPONGE * 2 + 1: adj3_t<2, NW_SPONGE>(g, io, shots, xm, st); break;
    case NW_CERJAN * 2 + 0: adj3_t<1, NW_CERJAN>(g, io, shots, xm, st); break;
    case NW_CERJAN * 2 + 1: adj3_t<2, NW_CERJAN>(g, io, shots, xm, st); break;
    default: break;
  }
}

}  // namespace nw
