// The generic leapfrog wave stepper and its adjoint: the C ABI of include/nefi_stepper.h.
//
// nefi/operators/timestepping.py:167-214 steps u_tt = c^2 lap(u) + s (optionally damped) with a
// second-order Laplacian whose boundary is a ghost cell per side (periodic wrap, replicated
// edge, zero) and, for the absorbing boundary, a first-order Mur update of the faces of each
// axis in turn (:145-164). Here one kernel does the Laplacian and the update per cell (the
// reference's operation order, contraction off, so float64 results match to rounding), and one
// small kernel per axis the Mur faces (it reads faces that earlier axes already updated, as the
// reference's sequence of concatenations does). The adjoint is written by hand: the Mur
// updates are undone per axis in reverse order, then one kernel applies the transposed
// interior update (the Laplacian with these ghosts is symmetric, so lap^T = lap).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "nefi_stepper.h"
#include "nefi_wave.h"
#include "nefi_wave_internal.h"

#pragma clang fp contract(off)

namespace nw {
namespace {

constexpr int kTsThreads = 256;

struct TsGeo {
  int64_t batch;
  int64_t n[3];   // extents
  int64_t st[3];  // element strides
  int64_t vol;    // cells of one grid
  int used[3];
  double h2[3];   // h^2 per axis
  double h[3];
  int bc;         // ghost rule of the Laplacian
};

TsGeo make_ts(const nw_ts_grid& g) {
  TsGeo t;
  t.batch = g.batch;
  for (int a = 0; a < 3; ++a) {
    t.n[a] = g.shape[a];
    t.used[a] = a >= 3 - g.ndim;
    t.h[a] = g.h[a];
    t.h2[a] = g.h[a] * g.h[a];
  }
  t.st[2] = 1;
  t.st[1] = t.n[2];
  t.st[0] = t.n[1] * t.n[2];
  t.vol = t.n[0] * t.n[1] * t.n[2];
  t.bc = g.boundary == NW_TS_ABSORBING ? NW_TS_NEUMANN : g.boundary;
  return t;
}

// sum over axes of ((u[-] - 2u) + u[+]) / h^2, starting from 0 and in axis order, with the
// ghost of each side given by the boundary rule (timestepping.py:60-87)
template <typename T, typename F>
__device__ __forceinline__ T ts_lap(const TsGeo& g, int64_t e, const int64_t* co, F f) {
  const T c = f(e);
  T out = T(0);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!g.used[a]) continue;
    const int64_t i = co[a], n = g.n[a], st = g.st[a];
    T m, p;
    if (i >= 1) {
      m = f(e - st);
    } else {
      m = g.bc == NW_TS_PERIODIC ? f(e + (n - 1) * st) : (g.bc == NW_TS_NEUMANN ? c : T(0));
    }
    if (i + 1 < n) {
      p = f(e + st);
    } else {
      p = g.bc == NW_TS_PERIODIC ? f(e - (n - 1) * st) : (g.bc == NW_TS_NEUMANN ? c : T(0));
    }
    out = out + ((m - T(2) * c) + p) / (T)g.h2[a];
  }
  return out;
}

__device__ __forceinline__ bool ts_cell(const TsGeo& g, int64_t e, int64_t* co) {
  if (e >= g.batch * g.vol) return false;
  const int64_t r = e % g.vol;
  co[0] = r / g.st[0];
  co[1] = (r / g.st[1]) % g.n[1];
  co[2] = r % g.n[2];
  return true;
}

template <typename T>
__global__ void __launch_bounds__(kTsThreads) k_ts_lap(TsGeo g, const T* u, T* out) {
  const int64_t e = (int64_t)blockIdx.x * kTsThreads + threadIdx.x;
  int64_t co[3];
  if (!ts_cell(g, e, co)) return;
  out[e] = ts_lap<T>(g, e, co, [&](int64_t j) { return u[j]; });
}

template <typename T>
__global__ void __launch_bounds__(kTsThreads) k_ts_step(TsGeo g, T dt, T dt2, const T* u_prev,
                                                        const T* u, const T* c, const T* s,
                                                        const T* gam, T* u_next) {
  const int64_t e = (int64_t)blockIdx.x * kTsThreads + threadIdx.x;
  int64_t co[3];
  if (!ts_cell(g, e, co)) return;
  const T lap = ts_lap<T>(g, e, co, [&](int64_t j) { return u[j]; });
  const T cv = c[e];
  T acc = (cv * cv) * lap;
  if (s) acc = acc + s[e];
  const T uc = u[e], up = u_prev[e];
  if (gam) {
    const T gd = (T(0.5) * gam[e]) * dt;
    u_next[e] = (((T(2) * uc) - ((T(1) - gd) * up)) + dt2 * acc) / (T(1) + gd);
  } else {
    u_next[e] = ((T(2) * uc) - up) + dt2 * acc;
  }
}

// one thread per line along `axis` (all other coordinates): both faces of the line
__device__ __forceinline__ bool ts_line(const TsGeo& g, int axis, int64_t t, int64_t& e0) {
  const int64_t lines = g.batch * (g.vol / g.n[axis]);
  if (t >= lines) return false;
  const int64_t b = t / (g.vol / g.n[axis]);
  int64_t r = t % (g.vol / g.n[axis]);
  int64_t co[3];
  for (int a = 2; a >= 0; --a) {
    if (a == axis) {
      co[a] = 0;
      continue;
    }
    co[a] = r % g.n[a];
    r /= g.n[a];
  }
  e0 = b * g.vol + co[0] * g.st[0] + co[1] * g.st[1] + co[2];
  return true;
}

template <typename T>
__device__ __forceinline__ T mur_k(T c, T dt, T h) {
  return (c * dt - h) / (c * dt + h);
}

template <typename T>
__global__ void __launch_bounds__(kTsThreads) k_ts_mur(TsGeo g, int axis, T dt, const T* u,
                                                       const T* c, T* w) {
  int64_t e0;
  if (!ts_line(g, axis, (int64_t)blockIdx.x * kTsThreads + threadIdx.x, e0)) return;
  const int64_t st = g.st[axis], n = g.n[axis];
  const T h = (T)g.h[axis];
  const int64_t e1 = e0 + st, en = e0 + (n - 1) * st, em = en - st;
  const T k_lo = mur_k(c[e0], dt, h), k_hi = mur_k(c[en], dt, h);
  const T lo = u[e1] + k_lo * (w[e1] - u[e0]);
  const T hi = u[em] + k_hi * (w[em] - u[en]);
  w[e0] = lo;
  w[en] = hi;
}

// dk/dc of k = (c dt - h)/(c dt + h)
template <typename T>
__device__ __forceinline__ T mur_dk(T c, T dt, T h) {
  const T den = c * dt + h;
  return (T(2) * h * dt) / (den * den);
}

template <typename T>
__global__ void __launch_bounds__(kTsThreads) k_ts_mur_adj(TsGeo g, int axis, T dt, const T* u,
                                                           const T* c, const T* w_prev, T* gw,
                                                           T* gu, T* gc) {
  int64_t e0;
  if (!ts_line(g, axis, (int64_t)blockIdx.x * kTsThreads + threadIdx.x, e0)) return;
  const int64_t st = g.st[axis], n = g.n[axis];
  const T h = (T)g.h[axis];
  const int64_t e1 = e0 + st, en = e0 + (n - 1) * st, em = en - st;
  const T k_lo = mur_k(c[e0], dt, h), k_hi = mur_k(c[en], dt, h);
  const T g_lo = gw[e0], g_hi = gw[en];
  gw[e0] = T(0);
  gw[en] = T(0);
  gw[e1] += k_lo * g_lo;  // (e1 == em when n == 3: one thread owns both updates)
  gw[em] += k_hi * g_hi;
  gu[e1] += g_lo;
  gu[e0] -= k_lo * g_lo;
  gu[em] += g_hi;
  gu[en] -= k_hi * g_hi;
  if (gc) {
    gc[e0] += g_lo * (w_prev[e1] - u[e0]) * mur_dk(c[e0], dt, h);
    gc[en] += g_hi * (w_prev[em] - u[en]) * mur_dk(c[en], dt, h);
  }
}

template <typename T>
__global__ void __launch_bounds__(kTsThreads) k_ts_step_adj(TsGeo g, T dt, T dt2, const T* u_prev,
                                                            const T* u, const T* c, const T* s,
                                                            const T* gam, const T* gw, T* gup,
                                                            T* gu, T* gc, T* gs, T* ggam) {
  const int64_t e = (int64_t)blockIdx.x * kTsThreads + threadIdx.x;
  int64_t co[3];
  if (!ts_cell(g, e, co)) return;
  // g = gw / (1 + gamma dt / 2) at any cell j
  auto gn = [&](int64_t j) {
    const T v = gw[j];
    return gam ? v / (T(1) + (T(0.5) * gam[j]) * dt) : v;
  };
  const T gne = gn(e);
  // u: 2 g + lap(q), q = dt^2 c^2 g (the Laplacian is symmetric)
  const T lq = ts_lap<T>(g, e, co, [&](int64_t j) {
    const T cj = c[j];
    return dt2 * ((cj * cj) * gn(j));
  });
  gu[e] += T(2) * gne + lq;
  const T gd = gam ? (T(0.5) * gam[e]) * dt : T(0);
  gup[e] = -((T(1) - gd) * gne);
  if (gc || ggam) {
    const T lap = ts_lap<T>(g, e, co, [&](int64_t j) { return u[j]; });
    const T cv = c[e];
    if (gc) gc[e] += T(2) * cv * lap * dt2 * gne;
    if (ggam) {
      T acc = (cv * cv) * lap;
      if (s) acc = acc + s[e];
      const T v = (((T(2) * u[e]) - ((T(1) - gd) * u_prev[e])) + dt2 * acc) / (T(1) + gd);
      ggam[e] = gne * (T(0.5) * dt) * (u_prev[e] - v);
    }
  }
  if (gs) gs[e] = dt2 * gne;
}

int check_grid(const nw_ts_grid* g) {
  if (!g) return fail(NW_ERR_ARG, "stepper: null grid");
  if (g->ndim < 1 || g->ndim > 3) return fail(NW_ERR_ARG, "stepper: ndim must be 1..3");
  if (g->dtype != NW_F32 && g->dtype != NW_F64) return fail(NW_ERR_ARG, "stepper: bad dtype");
  if (g->boundary < NW_TS_PERIODIC || g->boundary > NW_TS_ABSORBING)
    return fail(NW_ERR_ARG, "stepper: unknown boundary");
  if (g->batch < 0) return fail(NW_ERR_ARG, "stepper: batch must be >= 0");
  for (int a = 0; a < 3; ++a) {
    if (g->shape[a] < 1) return fail(NW_ERR_ARG, "stepper: extents must be >= 1");
    if (a < 3 - g->ndim && g->shape[a] != 1)
      return fail(NW_ERR_ARG, "stepper: unused leading extents must be 1");
    if (a >= 3 - g->ndim && !(g->h[a] > 0.0))
      return fail(NW_ERR_ARG, "stepper: spacing must be > 0");
  }
  return NW_OK;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kTsThreads - 1) / kTsThreads); }

int launched(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(NW_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return NW_OK;
}

}  // namespace
}  // namespace nw

using nw::fail;

extern "C" {

int32_t nw_ts_laplacian(const nw_ts_grid* g, const void* u, void* out, void* stream) {
  const int rc = nw::check_grid(g);
  if (rc != NW_OK) return rc;
  if (!u || !out) return fail(NW_ERR_ARG, "nw_ts_laplacian: null buffer");
  const nw::TsGeo t = nw::make_ts(*g);
  const int64_t cells = t.batch * t.vol;
  if (cells == 0) return NW_OK;
  hipStream_t st = (hipStream_t)stream;
  if (g->dtype == NW_F32)
    hipLaunchKernelGGL(nw::k_ts_lap<float>, dim3(nw::blocks_for(cells)), dim3(nw::kTsThreads), 0,
                       st, t, (const float*)u, (float*)out);
  else
    hipLaunchKernelGGL(nw::k_ts_lap<double>, dim3(nw::blocks_for(cells)), dim3(nw::kTsThreads), 0,
                       st, t, (const double*)u, (double*)out);
  return nw::launched("k_ts_lap");
}

int32_t nw_ts_mur(const nw_ts_grid* g, int32_t axis, double dt, const void* u, const void* c,
                  void* w, void* stream) {
  const int rc = nw::check_grid(g);
  if (rc != NW_OK) return rc;
  if (!u || !c || !w) return fail(NW_ERR_ARG, "nw_ts_mur: null buffer");
  if (axis < 3 - g->ndim || axis > 2 || g->shape[axis] < 3)
    return fail(NW_ERR_ARG, "nw_ts_mur: axis must be a grid axis of extent >= 3");
  const nw::TsGeo t = nw::make_ts(*g);
  const int64_t lines = t.batch * (t.vol / t.n[axis]);
  if (lines == 0) return NW_OK;
  hipStream_t st = (hipStream_t)stream;
  if (g->dtype == NW_F32)
    hipLaunchKernelGGL(nw::k_ts_mur<float>, dim3(nw::blocks_for(lines)), dim3(nw::kTsThreads), 0,
                       st, t, axis, (float)dt, (const float*)u, (const float*)c, (float*)w);
  else
    hipLaunchKernelGGL(nw::k_ts_mur<double>, dim3(nw::blocks_for(lines)), dim3(nw::kTsThreads), 0,
                       st, t, axis, dt, (const double*)u, (const double*)c, (double*)w);
  return nw::launched("k_ts_mur");
}

int32_t nw_ts_wave_step(const nw_ts_grid* g, double dt, const void* u_prev, const void* u,
                        const void* c, const void* s, const void* gamma, void* u_next,
                        int32_t mur, void* stream) {
  const int rc = nw::check_grid(g);
  if (rc != NW_OK) return rc;
  if (!u_prev || !u || !c || !u_next) return fail(NW_ERR_ARG, "nw_ts_wave_step: null buffer");
  const nw::TsGeo t = nw::make_ts(*g);
  const int64_t cells = t.batch * t.vol;
  if (cells == 0) return NW_OK;
  hipStream_t st = (hipStream_t)stream;
  const double dt2 = dt * dt;
  if (g->dtype == NW_F32)
    hipLaunchKernelGGL(nw::k_ts_step<float>, dim3(nw::blocks_for(cells)), dim3(nw::kTsThreads), 0,
                       st, t, (float)dt, (float)dt2, (const float*)u_prev, (const float*)u,
                       (const float*)c, (const float*)s, (const float*)gamma, (float*)u_next);
  else
    hipLaunchKernelGGL(nw::k_ts_step<double>, dim3(nw::blocks_for(cells)), dim3(nw::kTsThreads),
                       0, st, t, dt, dt2, (const double*)u_prev, (const double*)u,
                       (const double*)c, (const double*)s, (const double*)gamma,
                       (double*)u_next);
  int r = nw::launched("k_ts_step");
  if (r != NW_OK || !mur || g->boundary != NW_TS_ABSORBING) return r;
  for (int a = 3 - g->ndim; a < 3; ++a) {
    if (g->shape[a] < 3) continue;
    r = nw_ts_mur(g, a, dt, u, c, u_next, stream);
    if (r != NW_OK) return r;
  }
  return NW_OK;
}

int32_t nw_ts_mur_adjoint(const nw_ts_grid* g, int32_t axis, double dt, const void* u,
                          const void* c, const void* w_prev, void* g_w, void* g_u, void* g_c,
                          void* stream) {
  const int rc = nw::check_grid(g);
  if (rc != NW_OK) return rc;
  if (!u || !c || !g_w || !g_u || (g_c && !w_prev))
    return fail(NW_ERR_ARG, "nw_ts_mur_adjoint: null buffer");
  if (axis < 3 - g->ndim || axis > 2 || g->shape[axis] < 3)
    return fail(NW_ERR_ARG, "nw_ts_mur_adjoint: axis must be a grid axis of extent >= 3");
  const nw::TsGeo t = nw::make_ts(*g);
  const int64_t lines = t.batch * (t.vol / t.n[axis]);
  if (lines == 0) return NW_OK;
  hipStream_t st = (hipStream_t)stream;
  if (g->dtype == NW_F32)
    hipLaunchKernelGGL(nw::k_ts_mur_adj<float>, dim3(nw::blocks_for(lines)), dim3(nw::kTsThreads),
                       0, st, t, axis, (float)dt, (const float*)u, (const float*)c,
                       (const float*)w_prev, (float*)g_w, (float*)g_u, (float*)g_c);
  else
    hipLaunchKernelGGL(nw::k_ts_mur_adj<double>, dim3(nw::blocks_for(lines)),
                       dim3(nw::kTsThreads), 0, st, t, axis, dt, (const double*)u,
                       (const double*)c, (const double*)w_prev, (double*)g_w, (double*)g_u,
                       (double*)g_c);
  return nw::launched("k_ts_mur_adj");
}

int32_t nw_ts_wave_step_adjoint(const nw_ts_grid* g, double dt, const void* u_prev,
                                const void* u, const void* c, const void* s, const void* gamma,
                                const void* g_w, void* g_u_prev, void* g_u, void* g_c, void* g_s,
                                void* g_gamma, void* stream) {
  const int rc = nw::check_grid(g);
  if (rc != NW_OK) return rc;
  if (!u_prev || !u || !c || !g_w || !g_u_prev || !g_u)
    return fail(NW_ERR_ARG, "nw_ts_wave_step_adjoint: null buffer");
  if (g_gamma && !gamma) return fail(NW_ERR_ARG, "nw_ts_wave_step_adjoint: g_gamma needs gamma");
  const nw::TsGeo t = nw::make_ts(*g);
  const int64_t cells = t.batch * t.vol;
  if (cells == 0) return NW_OK;
  hipStream_t st = (hipStream_t)stream;
  const double dt2 = dt * dt;
  if (g->dtype == NW_F32)
    hipLaunchKernelGGL(nw::k_ts_step_adj<float>, dim3(nw::blocks_for(cells)),
                       dim3(nw::kTsThreads), 0, st, t, (float)dt, (float)dt2,
                       (const float*)u_prev, (const float*)u, (const float*)c, (const float*)s,
                       (const float*)gamma, (const float*)g_w, (float*)g_u_prev, (float*)g_u,
                       (float*)g_c, (float*)g_s, (float*)g_gamma);
  else
    hipLaunchKernelGGL(nw::k_ts_step_adj<double>, dim3(nw::blocks_for(cells)),
                       dim3(nw::kTsThreads), 0, st, t, dt, dt2, (const double*)u_prev,
                       (const double*)u, (const double*)c, (const double*)s,
                       (const double*)gamma, (const double*)g_w, (double*)g_u_prev,
                       (double*)g_u, (double*)g_c, (double*)g_s, (double*)g_gamma);
  return nw::launched("k_ts_step_adj");
}

}  // extern "C"
