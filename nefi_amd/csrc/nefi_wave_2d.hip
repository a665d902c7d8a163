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
__device__ __forceinline__ float4 row4(const float* g, int y, int D1, int D2, int x, bool act) {
  return (act && y >= 0 && y < D1) ? ld4(g + (int64_t)y * D2 + x) : f4(0.f);
}

// x-neighbours of a row vector v held by every lane at columns x..x+3: returns the values at
// x-2, x-1 (l2, l1) and x+4, x+5 (r1, r2). Lanes 0 and 63 read their outer columns from `g`.
__device__ __forceinline__ void xhalo(const float4& v, const float* g, int y, int D2, int x,
                                      int lane, float& l2, float& l1, float& r1, float& r2) {
  l2 = __shfl_up(v.z, 1, 64);
  l1 = __shfl_up(v.w, 1, 64);
  r1 = __shfl_down(v.x, 1, 64);
  r2 = __shfl_down(v.y, 1, 64);
  if (lane == 0) {
    const float* rp = g + (int64_t)y * D2;
    l2 = x - 2 >= 0 ? rp[x - 2] : 0.f;
    l1 = x - 1 >= 0 ? rp[x - 1] : 0.f;
  }
  if (lane == 63) {
    const float* rp = g + (int64_t)y * D2;
    r1 = x + 4 < D2 ? rp[x + 4] : 0.f;
    r2 = x + 5 < D2 ? rp[x + 5] : 0.f;
  }
}

// same for a product of two grids (q = a*b), used by the reverse sweep for m*a_eff*lambda
__device__ __forceinline__ void xhalo_prod(const float4& v, const float* ga, const float* gb,
                                           int y, int D2, int x, int lane, float& l2, float& l1,
                                           float& r1, float& r2) {
  l2 = __shfl_up(v.z, 1, 64);
  l1 = __shfl_up(v.w, 1, 64);
  r1 = __shfl_down(v.x, 1, 64);
  r2 = __shfl_down(v.y, 1, 64);
  if (lane == 0) {
    const int64_t o = (int64_t)y * D2;
    l2 = x - 2 >= 0 ? ga[o + x - 2] * gb[o + x - 2] : 0.f;
    l1 = x - 1 >= 0 ? ga[o + x - 1] * gb[o + x - 1] : 0.f;
  }
  if (lane == 63) {
    const int64_t o = (int64_t)y * D2;
    r1 = x + 4 < D2 ? ga[o + x + 4] * gb[o + x + 4] : 0.f;
    r2 = x + 5 < D2 ? ga[o + x + 5] * gb[o + x + 5] : 0.f;
  }
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
// forward step
// ------------------------------------------------------------------------------------------
template <int R, int ABS, int TY>
__global__ void __launch_bounds__(256) k2_step(Geo<float> g, StepIO<float> io, int shots,
                                               int spw) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2];
  const int x = (blockIdx.x * 64 + lane) * 4;
  const int ys = (blockIdx.y * 4 + wave) * TY;
  if (ys >= D1) return;
  const int ye = ys + TY < D1 ? ys + TY : D1;
  const bool act = x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const float w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.f;
  const int s_end = min(shots, (int)(blockIdx.z + 1) * spw);
  for (int s = blockIdx.z * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const float* P = io.p + base;
    float4 w[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) w[k] = row4(P, ys - R + k, D1, D2, x, act);
#pragma unroll 2
    for (int y = ys; y < ye; ++y) {
      w[2 * R] = row4(P, y + R, D1, D2, x, act);
      const int64_t ro = (int64_t)y * D2 + x;
      float l2, l1, r1, r2;
      xhalo(w[R], P, y, D2, x, lane, l2, l1, r1, r2);
      if (act) {
        const float4 lap = stencil4<R>(g, w, l2, l1, r1, r2);
        const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
        const float4 pp = ld4(io.p_prev + base + ro);
        const float4 mm = ld4(io.m + ro);
        const float4 pc = w[R];
        float4 out;
        float o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          float a, b, tap;
          damp1<ABS>(hy, comp(hx, k), a, b, tap);
          const float prev_raw = comp(pp, k);
          const float prev = (ABS == NW_CERJAN && io.taper_prev) ? tap * prev_raw : prev_raw;
          const float rhs = (2.f * comp(pc, k) - b * prev) + comp(mm, k) * comp(lap, k);
          o[k] = rhs;
        }
        if (io.src_taps > 0) {
          for (int t = 0; t < io.src_taps; ++t) {
            const int64_t d = (int64_t)io.src_idx[s * io.src_taps + t] - ro;
            if (d >= 0 && d < 4) {
              const float v = io.src_val[s * io.src_taps + t] * w_n;
              if (d == 0) o[0] += v;
              else if (d == 1) o[1] += v;
              else if (d == 2) o[2] += v;
              else o[3] += v;
            }
          }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          float a, b, tap;
          damp1<ABS>(hy, comp(hx, k), a, b, tap);
          if (ABS != NW_NONE) o[k] = a * o[k];
        }
        out = make_float4(o[0], o[1], o[2], o[3]);
        st4(io.p_next + base + ro, out);
        if (io.ck_prev) {
          st4(io.ck_prev + base + ro, pp);
          st4(io.ck_cur + base + ro, pc);
        }
      }
#pragma unroll
      for (int k = 0; k < 2 * R; ++k) w[k] = w[k + 1];
    }
  }
}

// ------------------------------------------------------------------------------------------
// reverse step with the imaging sum
// ------------------------------------------------------------------------------------------
template <int R, int ABS, int TY>
__global__ void __launch_bounds__(256) k2_adjoint(Geo<float> g, AdjIO<float> io, int shots,
                                                  int spw) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int D1 = (int)g.D[1], D2 = (int)g.D[2];
  const int x = (blockIdx.x * 64 + lane) * 4;
  const int ys = (blockIdx.y * 4 + wave) * TY;
  if (ys >= D1) return;
  const int ye = ys + TY < D1 ? ys + TY : D1;
  const bool act = x < D2;
  const float4 hx = (ABS != NW_NONE && act) ? ld4(g.damp[2] + x) : f4(0.f);
  const int s_end = min(shots, (int)(blockIdx.z + 1) * spw);
  // gradient slice of this shot group (one float64 accumulator grid per group)
  double* G = io.grad ? io.grad + (int64_t)blockIdx.z * g.cells : nullptr;
  for (int s = blockIdx.z * spw; s < s_end; ++s) {
    const int64_t base = (int64_t)s * g.cells;
    const float* L1 = io.lam1 + base;
    const float* MA = io.ma;
    const float* Pn = io.p_n ? io.p_n + base : nullptr;
    float4 lw[2 * R + 1], mw[2 * R + 1], pw[2 * R + 1];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) {
      lw[k] = row4(L1, ys - R + k, D1, D2, x, act);
      mw[k] = row4(MA, ys - R + k, D1, D2, x, act);
      pw[k] = Pn ? row4(Pn, ys - R + k, D1, D2, x, act) : f4(0.f);
    }
#pragma unroll 2
    for (int y = ys; y < ye; ++y) {
      lw[2 * R] = row4(L1, y + R, D1, D2, x, act);
      mw[2 * R] = row4(MA, y + R, D1, D2, x, act);
      if (Pn) pw[2 * R] = row4(Pn, y + R, D1, D2, x, act);
      const int64_t ro = (int64_t)y * D2 + x;
      float4 qw[2 * R + 1];
#pragma unroll
      for (int k = 0; k < 2 * R + 1; ++k)
        qw[k] = make_float4(mw[k].x * lw[k].x, mw[k].y * lw[k].y, mw[k].z * lw[k].z,
                            mw[k].w * lw[k].w);
      float l2, l1, r1, r2;
      xhalo_prod(qw[R], MA, L1, y, D2, x, lane, l2, l1, r1, r2);
      float pl2 = 0.f, pl1 = 0.f, pr1 = 0.f, pr2 = 0.f;
      if (Pn) xhalo(pw[R], Pn, y, D2, x, lane, pl2, pl1, pr1, pr2);
      if (act) {
        const float4 lapq = stencil4<R>(g, qw, l2, l1, r1, r2);
        const float hy = ABS != NW_NONE ? g.damp[1][y] : 0.f;
        const float4 l2c = ld4(io.lam2 + base + ro);
        const float4 lc = lw[R];
        float o[4], rb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          float a, b, tap;
          damp1<ABS>(hy, comp(hx, k), a, b, tap);
          const float rbar = a * comp(lc, k);
          const float tp1 = (ABS == NW_CERJAN && io.taper_next) ? tap : 1.f;
          const float pbar_next = -(b * tp1) * (a * comp(l2c, k));
          o[k] = pbar_next + 2.f * rbar + comp(lapq, k);
          rb[k] = rbar;
        }
        st4(io.lam0 + base + ro, make_float4(o[0], o[1], o[2], o[3]));
        if (Pn && G) {
          const float4 lapp = stencil4<R>(g, pw, pl2, pl1, pr1, pr2);
          double2* gp = reinterpret_cast<double2*>(G + ro);
          double2 g0 = gp[0], g1 = gp[1];
          g0.x += (double)rb[0] * (double)lapp.x;
          g0.y += (double)rb[1] * (double)lapp.y;
          g1.x += (double)rb[2] * (double)lapp.z;
          g1.y += (double)rb[3] * (double)lapp.w;
          gp[0] = g0;
          gp[1] = g1;
        }
      }
#pragma unroll
      for (int k = 0; k < 2 * R; ++k) {
        lw[k] = lw[k + 1];
        mw[k] = mw[k + 1];
        pw[k] = pw[k + 1];
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

template <int R, int ABS, int TY>
static void launch_step_t(const Geo<float>& g, const StepIO<float>& io, int shots, int spw,
                          hipStream_t st) {
  const dim3 grid((unsigned)((g.D[2] + 255) / 256), (unsigned)((g.D[1] + 4 * TY - 1) / (4 * TY)),
                  (unsigned)((shots + spw - 1) / spw));
  hipLaunchKernelGGL((k2_step<R, ABS, TY>), grid, dim3(256), 0, st, g, io, shots, spw);
}

template <int R, int ABS, int TY>
static void launch_adj_t(const Geo<float>& g, const AdjIO<float>& io, int shots, int spw,
                         hipStream_t st) {
  const dim3 grid((unsigned)((g.D[2] + 255) / 256), (unsigned)((g.D[1] + 4 * TY - 1) / (4 * TY)),
                  (unsigned)((shots + spw - 1) / spw));
  hipLaunchKernelGGL((k2_adjoint<R, ABS, TY>), grid, dim3(256), 0, st, g, io, shots, spw);
}

template <int R, int ABS>
static void step_ty(const Geo<float>& g, const StepIO<float>& io, int shots, int ty, int spw,
                    hipStream_t st) {
  switch (ty) {
    case 4: launch_step_t<R, ABS, 4>(g, io, shots, spw, st); break;
    case 8: launch_step_t<R, ABS, 8>(g, io, shots, spw, st); break;
    case 32: launch_step_t<R, ABS, 32>(g, io, shots, spw, st); break;
    case 64: launch_step_t<R, ABS, 64>(g, io, shots, spw, st); break;
    default: launch_step_t<R, ABS, 16>(g, io, shots, spw, st); break;
  }
}

template <int R, int ABS>
static void adj_ty(const Geo<float>& g, const AdjIO<float>& io, int shots, int ty, int spw,
                   hipStream_t st) {
  switch (ty) {
    case 4: launch_adj_t<R, ABS, 4>(g, io, shots, spw, st); break;
    case 8: launch_adj_t<R, ABS, 8>(g, io, shots, spw, st); break;
    case 32: launch_adj_t<R, ABS, 32>(g, io, shots, spw, st); break;
    case 64: launch_adj_t<R, ABS, 64>(g, io, shots, spw, st); break;
    default: launch_adj_t<R, ABS, 16>(g, io, shots, spw, st); break;
  }
}

void fast2d_step(const Geo<float>& g, int order, int absorbing, const StepIO<float>& io,
                 int shots, int ty, int spw, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int key = absorbing * 2 + (order == 4 ? 1 : 0);
  switch (key) {
    case NW_NONE * 2 + 0: step_ty<1, NW_NONE>(g, io, shots, ty, spw, st); break;
    case NW_NONE * 2 + 1: step_ty<2, NW_NONE>(g, io, shots, ty, spw, st); break;
    case NW_SPONGE * 2 + 0: step_ty<1, NW_SPONGE>(g, io, shots, ty, spw, st); break;
    case NW_SPONGE * 2 + 1: step_ty<2, NW_SPONGE>(g, io, shots, ty, spw, st); break;
    case NW_CERJAN * 2 + 0: step_ty<1, NW_CERJAN>(g, io, shots, ty, spw, st); break;
    default: step_ty<2, NW_CERJAN>(g, io, shots, ty, spw, st); break;
  }
}

void fast2d_adjoint(const Geo<float>& g, int order, int absorbing, const AdjIO<float>& io,
                    int shots, int ty, int spw, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int key = absorbing * 2 + (order == 4 ? 1 : 0);
  switch (key) {
    case NW_NONE * 2 + 0: adj_ty<1, NW_NONE>(g, io, shots, ty, spw, st); break;
    case NW_NONE * 2 + 1: adj_ty<2, NW_NONE>(g, io, shots, ty, spw, st); break;
    case NW_SPONGE * 2 + 0: adj_ty<1, NW_SPONGE>(g, io, shots, ty, spw, st); break;
    case NW_SPONGE * 2 + 1: adj_ty<2, NW_SPONGE>(g, io, shots, ty, spw, st); break;
    case NW_CERJAN * 2 + 0: adj_ty<1, NW_CERJAN>(g, io, shots, ty, spw, st); break;
    default: adj_ty<2, NW_CERJAN>(g, io, shots, ty, spw, st); break;
  }
}

}  // namespace nw
