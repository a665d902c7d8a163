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
// Arithmetic follows the generic kernels (nefi_wave.hip) operation by operation.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nefi_wave.h"
#include "nefi_wave_internal.h"

namespace nw {

namespace {

__device__ __forceinline__ float4 ld4(const float* p) {
  return *reinterpret_cast<const float4*>(p);
}
__device__ __forceinline__ void st4(float* p, const float4& v) {
  *reinterpret_cast<float4*>(p) = v;
}
__device__ __forceinline__ float4 f4(float a) { return make_float4(a, a, a, a); }
__device__ __forceinline__ float comp(const float4& v, int k) {
  return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
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
  float s = g.cc * c;
  float ay = g.c1[1] * (u1 + d1);
  if (R == 2) ay += g.c2[1] * (u2 + d2);
  s += ay;
  float ax = g.c1[2] * (l1 + r1);
  if (R == 2) ax += g.c2[2] * (l2 + r2);
  s += ax;
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
  if (ABS == NW_SPONGE) {
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
  return (2.f * pc - b * prev) + m * lap;
}

// adjoint of one cell: lambda^n from rbar = a lambda^{n+1}, the next step's lambda^{n+2} term
// and the stencil of q
template <int ABS>
__device__ __forceinline__ float adj1(float lc, float l2c, float lapq, float a, float b,
                                      float tap, int taper_next, float& rbar) {
  rbar = a * lc;
  const float tp1 = (ABS == NW_CERJAN && taper_next) ? tap : 1.f;
  const float pbar_next = -(b * tp1) * (a * l2c);
  return pbar_next + 2.f * rbar + lapq;
}

__device__ __forceinline__ float4 mul4(const float4& a, const float4& b) {
  return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w);
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
    const int d = io.src_idx[s * io.src_taps + t] - ro;
    if (d >= 0 && d < 4) {
      const float v = io.src_val[s * io.src_taps + t] * w_n;
      o.x += d == 0 ? v : 0.f;
      o.y += d == 1 ? v : 0.f;
      o.z += d == 2 ? v : 0.f;
      o.w += d == 3 ? v : 0.f;
    }
  }
  if (ABS != NW_NONE) o = mul4(dm.a, o);
  st4(io.p_next + base + ro, o);
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
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
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
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
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
      const float4 l2c = act ? ld4(io.lam2 + base + ro) : f4(0.f);
      const float4 lc = act ? ld4(L1 + ro) : f4(0.f);  // read R rows ago: a cache hit
      const float2 el = edge2(L1, y, D2, P2, x, lane);
      const float2 em = edge2(MA, y, D2, P2, x, lane);
      const float2 ep = Pn ? edge2(Pn, y, D2, P2, x, lane) : make_float2(0.f, 0.f);
      double2* gp = (G && Pn && act) ? reinterpret_cast<double2*>(G + ro) : nullptr;
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
          const float4 lapp = stencil4<R>(g, pw, pnb.x, pnb.y, pnb.z, pnb.w);
          g0.x += (double)rb.x * (double)lapp.x;
          g0.y += (double)rb.y * (double)lapp.y;
          g1.x += (double)rb.z * (double)lapp.z;
          g1.y += (double)rb.w * (double)lapp.w;
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

template <int TY, int R>
struct TileShape {
  static constexpr int rows = TY + 2 * R;
  static constexpr int n4 = rows * (kLW / 4);  // float4 elements of one tile
  static constexpr int per_thread = (n4 + 255) / 256;
  static constexpr int my_rows = TY / 4;
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
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  const int x = x0 + 4 * lane;
  const int r0 = wave * S::my_rows;  // first tile row of this wavefront
  const bool act = x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const float w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.f;
  // model rows of this wavefront: shot-independent, loaded once for the group
  float4 mm[S::my_rows];
#pragma unroll
  for (int k = 0; k < S::my_rows; ++k) {
    const int y = y0 + r0 + k;
    mm[k] = (act && y < D1) ? ld4(io.m + (y * P2 + x)) : f4(0.f);
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
      pp[k] = (act && y < D1) ? ld4(io.p_prev + base + (y * P2 + x)) : f4(0.f);
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
        const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
        const Damp4<ABS> dm = damp4<ABS>(hy, hx);
        const float4 pc = w[R];
        const int tpv = io.taper_prev;
        float4 o = make_float4(rhs1<ABS>(pc.x, pp[k].x, mm[k].x, lap.x, dm.b.x, dm.tap.x, tpv),
                               rhs1<ABS>(pc.y, pp[k].y, mm[k].y, lap.y, dm.b.y, dm.tap.y, tpv),
                               rhs1<ABS>(pc.z, pp[k].z, mm[k].z, lap.z, dm.b.z, dm.tap.z, tpv),
                               rhs1<ABS>(pc.w, pp[k].w, mm[k].w, lap.w, dm.b.w, dm.tap.w, tpv));
        for (int t = 0; t < io.src_taps; ++t) {
          const int d = io.src_idx[s * io.src_taps + t] - ro;
          if (d >= 0 && d < 4) {
            const float sv = io.src_val[s * io.src_taps + t] * w_n;
            o.x += d == 0 ? sv : 0.f;
            o.y += d == 1 ? sv : 0.f;
            o.z += d == 2 ? sv : 0.f;
            o.w += d == 3 ? sv : 0.f;
          }
        }
        if (ABS != NW_NONE) o = mul4(dm.a, o);
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

template <int R, int ABS, int TY>
__global__ void __launch_bounds__(256) k3_adjoint(Geo<float> g, AdjIO<float> io, int shots,
                                                  int spw, int xcd_map) {
  using S = TileShape<TY, R>;
  __shared__ __align__(16) float lq[S::rows * kLW];  // q = (m a_eff) lambda^{n+1}
  __shared__ __align__(16) float lp[S::rows * kLW];  // p^n
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const Tile tl = tile_of(div_up(D2, 256), div_up(D1, TY), div_up(shots, spw), xcd_map);
  if (!tl.valid) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x0 = tl.cx * 256, y0 = tl.ry * TY;
  const int x = x0 + 4 * lane;
  const int r0 = wave * S::my_rows;
  const bool act = x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const bool imaging = io.p_n != nullptr && io.grad != nullptr;
  const int tpn = io.taper_next;
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
    {
      float4 vl[S::per_thread], vm[S::per_thread];
      tile_load<TY, R>(L1, y0, x0, D1, D2, P2, vl);
      tile_load<TY, R>(io.ma, y0, x0, D1, D2, P2, vm);
      __syncthreads();  // the previous shot's reads of lq / lp are done
#pragma unroll
      for (int k = 0; k < S::per_thread; ++k) vl[k] = mul4(vm[k], vl[k]);
      tile_store<TY, R>(lq, vl);
    }
    if (imaging) {
      float4 vp[S::per_thread];
      tile_load<TY, R>(io.p_n + base, y0, x0, D1, D2, P2, vp);
      tile_store<TY, R>(lp, vp);
    }
    float4 lc[S::my_rows], l2c[S::my_rows];
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      const bool ok = act && y < D1;
      lc[k] = ok ? ld4(L1 + (y * P2 + x)) : f4(0.f);  // just loaded into the tile: cache hit
      l2c[k] = ok ? ld4(io.lam2 + base + (y * P2 + x)) : f4(0.f);
    }
    __syncthreads();
    float4 qw[2 * R + 1], pw[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) {
      qw[k] = lrow(lq, r0 + k, lane);
      pw[k] = imaging ? lrow(lp, r0 + k, lane) : f4(0.f);
    }
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + r0 + k;
      qw[2 * R] = lrow(lq, r0 + k + 2 * R, lane);
      pw[2 * R] = imaging ? lrow(lp, r0 + k + 2 * R, lane) : f4(0.f);
      if (act && y < D1) {
        const float4 qn = lnb(lq, r0 + k + R, lane);
        const float4 lapq = stencil4<R>(g, qw, qn.x, qn.y, qn.z, qn.w);
        const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
        const Damp4<ABS> dm = damp4<ABS>(hy, hx);
        float4 rb;
        const float4 o = make_float4(
            adj1<ABS>(lc[k].x, l2c[k].x, lapq.x, dm.a.x, dm.b.x, dm.tap.x, tpn, rb.x),
            adj1<ABS>(lc[k].y, l2c[k].y, lapq.y, dm.a.y, dm.b.y, dm.tap.y, tpn, rb.y),
            adj1<ABS>(lc[k].z, l2c[k].z, lapq.z, dm.a.z, dm.b.z, dm.tap.z, tpn, rb.z),
            adj1<ABS>(lc[k].w, l2c[k].w, lapq.w, dm.a.w, dm.b.w, dm.tap.w, tpn, rb.w));
        st4(io.lam0 + base + (y * P2 + x), o);
        if (imaging) {
          const float4 pn = lnb(lp, r0 + k + R, lane);
          const float4 lapp = stencil4<R>(g, pw, pn.x, pn.y, pn.z, pn.w);
          ga[k].x += (double)rb.x * (double)lapp.x;
          ga[k].y += (double)rb.y * (double)lapp.y;
          gb[k].x += (double)rb.z * (double)lapp.z;
          gb[k].y += (double)rb.w * (double)lapp.w;
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

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
bool fast2d_supported(const Geo<float>& g, int absorbing, int dtype) {
  return g.nd == 2 && dtype == NW_F32 && absorbing != NW_PML && g.D[2] % 4 == 0;
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

template <int R, int ABS, int TY>
static void launch_step_t(const Geo<float>& g, const StepIO<float>& io, int shots, int spw,
                          int xm, hipStream_t st) {
  // xm bit 2 selects the LDS-tiled kernel, bit 1 the two-rows-per-iteration march
  if ((xm & 4) && TY >= 8 && TY <= 32)
    hipLaunchKernelGGL((k3_step<R, ABS, (TY < 8 ? 8 : (TY > 32 ? 32 : TY))>),
                       dim3(grid_tiles3(g, TY, shots, spw)), dim3(256), 0, st, g, io, shots, spw,
                       xm & 1);
  else if (xm & 2)
    hipLaunchKernelGGL((k2_step<R, ABS, TY, 2>), dim3(grid_tiles(g, TY, shots, spw)), dim3(256),
                       0, st, g, io, shots, spw, xm & 1);
  else
    hipLaunchKernelGGL((k2_step<R, ABS, TY, 1>), dim3(grid_tiles(g, TY, shots, spw)), dim3(256),
                       0, st, g, io, shots, spw, xm & 1);
}

template <int R, int ABS, int TY>
static void launch_adj_t(const Geo<float>& g, const AdjIO<float>& io, int shots, int spw, int xm,
                         hipStream_t st) {
  if ((xm & 4) && TY >= 8 && TY <= 32)
    hipLaunchKernelGGL((k3_adjoint<R, ABS, (TY < 8 ? 8 : (TY > 32 ? 32 : TY))>),
                       dim3(grid_tiles3(g, TY, shots, spw)), dim3(256), 0, st, g, io, shots, spw,
                       xm & 1);
  else
    hipLaunchKernelGGL((k2_adjoint<R, ABS, TY>), dim3(grid_tiles(g, TY, shots, spw)), dim3(256),
                       0, st, g, io, shots, spw, xm & 1);
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
    default: step_ty<2, NW_CERJAN>(g, io, shots, ty, spw, xm, st); break;
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
    default: adj_ty<2, NW_CERJAN>(g, io, shots, ty, spw, xm, st); break;
  }
}

}  // namespace nw
