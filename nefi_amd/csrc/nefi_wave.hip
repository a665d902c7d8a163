// libnefi_wave: acoustic leapfrog propagator for MI355X (gfx950).
//
// Generic kernel family: any of 1-3 dimensions, Laplacian order 2 or 4, the four absorbing
// layers of nefi (none, sponge, Cerjan taper, second-order PML), float or double. One thread
// owns one padded cell of one shot and reads its stencil neighbours straight from global
// memory (they are L1/L2 hits). The fast 2-D path lives in nefi_wave_2d.hip and is selected by
// the host loop for the shapes it supports.
//
// The step follows nefi/physics/wave.py:529-552 term by term:
//   lap  = L p^n (+ PML: sum_a (G_a(psi'_a + psi_a)) / (2 h_a),  psi'_a = alpha psi_a + beta D_a p^n)
//   rhs  = (2 p^n - b p^{n-1}) + m * lap  (- zz p^n in 2-D PML)  + src * w_n
//   p^{n+1} = a * rhs                      (Cerjan: taper * rhs, and the stored p^n is tapered
//                                           when it is read back as p^{n-1})
// and the reverse sweep is its exact transpose (see k_adjoint).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "nefi_wave.h"
#include "nefi_wave_internal.h"

namespace nw {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// ------------------------------------------------------------------------------------------
// optional per-launch timing (nw_profile_enable / nw_profile_read)
// ------------------------------------------------------------------------------------------
struct ProfRec {
  int kind;
  int64_t units;  // time steps (or samples) the launch covers
  hipEvent_t start, stop;
};

struct Profiler {
  bool on = false;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  void recycle() {
    for (auto& r : recs) {
      pool.push_back(r.start);
      pool.push_back(r.stop);
    }
    recs.clear();
  }
};

static Profiler g_prof;

// One captured call: its argument key, the instantiated graph and the pointer arrangement the
// call leaves behind.
struct GraphEntry {
  std::vector<uintptr_t> key;
  hipGraphExec_t exec = nullptr;
  std::vector<void*> out;
  hipEvent_t done = nullptr;  // recorded after the entry's latest launch
  int64_t last_use = 0;       // plan's launch clock at that launch
};

// brackets one launch; returns the record index or -1 when timing is off
int prof_begin(int kind, hipStream_t st, int64_t units) {
  if (!g_prof.on) return -1;
  ProfRec r{kind, units, g_prof.get(), g_prof.get()};
  if (!r.start || !r.stop) return -1;
  (void)hipEventRecord(r.start, st);  // a failure surfaces when nw_profile_read waits on it
  g_prof.recs.push_back(r);
  return (int)g_prof.recs.size() - 1;
}

void prof_end(int id, hipStream_t st) {
  if (id >= 0) (void)hipEventRecord(g_prof.recs[id].stop, st);
}

// ------------------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int64_t stride_of(const Geo<T>& g, int a) {
  return a == 0 ? g.D[1] * g.P2 : (a == 1 ? g.P2 : 1);
}

// flat index of face f along axis a at the other coordinates of `co` (rows of the axis 2 face
// grid are FP2 elements apart, so they start 16-byte aligned like the wavefield rows)
template <typename T>
__device__ __forceinline__ int64_t face_at(const Geo<T>& g, int a, const int64_t* co, int64_t f) {
  const int64_t d1 = g.D[1] + (a == 1), d2 = a == 2 ? g.FP2 : g.D[2];
  const int64_t c0 = a == 0 ? f : co[0], c1 = a == 1 ? f : co[1], c2 = a == 2 ? f : co[2];
  return (c0 * d1 + c1) * d2 + c2;
}

// damping factors of one cell: a_eff (applied to rhs), b (applied to p^{n-1}), taper
template <typename T, int ABS>
__device__ __forceinline__ void damping(const Geo<T>& g, const int64_t* co, T& a, T& b, T& tap) {
  a = T(1);
  b = T(1);
  tap = T(1);
  if (ABS == NW_SPONGE || ABS == NW_PML) {
    T half = T(0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      if (g.damp[ax]) half += g.damp[ax][co[ax]];
    a = T(1) / (T(1) + half);
    b = T(1) - half;
  } else if (ABS == NW_CERJAN) {
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      if (g.damp[ax]) {
        const T v = g.damp[ax][co[ax]];
        tap = v < tap ? v : tap;
      }
    a = tap;
  }
}

// stencil of a cell-valued function f(j) (j = flat index in the shot grid), zero outside
template <typename T, int R, typename F>
__device__ __forceinline__ T lap_of(const Geo<T>& g, const int64_t* co, int64_t idx, F f) {
  T s = g.cc * f(idx);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!g.used[a]) continue;
    const int64_t st = stride_of(g, a);
    const int64_t i = co[a], n = g.D[a];
    const T m1 = i >= 1 ? f(idx - st) : T(0);
    const T p1 = i + 1 < n ? f(idx + st) : T(0);
    T acc = g.c1[a] * (m1 + p1);
    if (R == 2) {
      const T m2 = i >= 2 ? f(idx - 2 * st) : T(0);
      const T p2 = i + 2 < n ? f(idx + 2 * st) : T(0);
      acc += g.c2[a] * (m2 + p2);
    }
    s += acc;
  }
  return s;
}

template <typename T>
__device__ __forceinline__ bool cell_of(const Geo<T>& g, int64_t& s, int64_t* co, int64_t& idx) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t y = (int64_t)blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= g.D[2] || y >= g.D[1]) return false;
  const int64_t zs = blockIdx.z;
  co[0] = zs % g.D[0];
  s = zs / g.D[0];
  co[1] = y;
  co[2] = x;
  idx = (co[0] * g.D[1] + y) * g.P2 + x;
  return true;
}

// ------------------------------------------------------------------------------------------
// forward step
// ------------------------------------------------------------------------------------------
template <typename T, int R, int ABS>
__global__ void __launch_bounds__(256) k_step(Geo<T> g, StepIO<T> io) {
  int64_t s, co[3], idx;
  if (!cell_of(g, s, co, idx)) return;
  const int64_t base = s * g.cells;
  const T* __restrict__ P = io.p + base;
  const T pc = P[idx];
  T lap = lap_of<T, R>(g, co, idx, [&](int64_t j) { return P[j]; });

  if (ABS == NW_PML) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!g.used[a]) continue;
      const int64_t st = stride_of(g, a), i = co[a], n = g.D[a];
      const T pm = i >= 1 ? P[idx - st] : T(0);
      const T pp = i + 1 < n ? P[idx + st] : T(0);
      T other = T(0);
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if (b != a && g.used[b]) other += g.zc[b][co[b]];
      const int64_t f0 = face_at(g, a, co, i), f1 = face_at(g, a, co, i + 1);
      const int64_t fb = s * g.fcells[a];
      const T o0 = io.psi_old[a][fb + f0], o1 = io.psi_old[a][fb + f1];
      const T b0 = (other - g.zf[a][i]) * g.kappa[a][i];
      const T b1 = (other - g.zf[a][i + 1]) * g.kappa[a][i + 1];
      const T n0 = g.alpha[a][i] * o0 + b0 * (pc - pm);
      const T n1 = g.alpha[a][i + 1] * o1 + b1 * (pp - pc);
      lap += ((n1 + o1) - (n0 + o0)) * g.inv2h[a];
      io.psi_new[a][fb + f0] = n0;
      if (i == n - 1) io.psi_new[a][fb + f1] = n1;
      if (io.ck_psi[a]) {
        io.ck_psi[a][fb + f0] = o0;
        if (i == n - 1) io.ck_psi[a][fb + f1] = o1;
      }
    }
  }

  T a_eff, b, tap;
  damping<T, ABS>(g, co, a_eff, b, tap);
  const T prev_raw = io.p_prev[base + idx];
  const T prev = (ABS == NW_CERJAN && io.taper_prev) ? tap * prev_raw : prev_raw;
  T rhs = (T(2) * pc - b * prev) + io.m[idx] * lap;
  if (ABS == NW_PML && g.nd == 2) rhs -= g.dt2 * g.zc[1][co[1]] * g.zc[2][co[2]] * pc;
  if (io.src_taps > 0) {
    const T w = io.wavelet[io.n];
    for (int k = 0; k < io.src_taps; ++k)
      if (io.src_idx[s * io.src_taps + k] == idx) rhs += io.src_val[s * io.src_taps + k] * w;
  }
  io.p_next[base + idx] = (ABS == NW_NONE) ? rhs : a_eff * rhs;
  if (io.ck_prev) {
    io.ck_prev[base + idx] = prev_raw;
    io.ck_cur[base + idx] = pc;
  }
}

// receiver gather of p^n (wave.py:321-324): groups of `taps` lanes hold one receiver's taps
// and are reduced with wavefront shuffles; lane 0 of the group writes the sample.
template <typename T>
__global__ void __launch_bounds__(256) k_sample(Geo<T> g, SampleIO<T> io) {
  const int s = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int taps = io.taps;
  const int r = t / taps, k = t % taps;
  const int count = io.rec_count[s];
  T v = T(0);
  if (r < count) {
    const int64_t row = (int64_t)io.rec_first[s] + r;
    v = io.rec_w[row * taps + k] * io.p[(int64_t)s * g.cells + io.rec_idx[row * taps + k]];
  }
  for (int off = 1; off < taps; off <<= 1) v += __shfl_xor(v, off, 64);
  if (r < count && k == 0)
    io.traces[((int64_t)io.trace_first[s] + r) * io.n_t + io.j] = v;
}

// lam^n += R^T g_n for one recorded step (deterministic: one thread per distinct cell)
template <typename T>
__device__ __forceinline__ T inject_sum(const InjectIO<T>& io, int s, int uu, int j) {
  const int64_t row0 = io.trace_first[s];
  T acc = T(0);
  for (int e = io.inj_off[uu]; e < io.inj_off[uu + 1]; ++e)
    acc += io.inj_w[e] * io.residual[(row0 + io.inj_rec[e]) * io.n_t + j];
  return acc;
}

template <typename T>
__global__ void __launch_bounds__(256) k_inject(Geo<T> g, InjectIO<T> io) {
  const int s = blockIdx.y;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= io.inj_count[s]) return;
  const int uu = io.inj_first[s] + u;
  io.lam[(int64_t)s * g.cells + io.inj_cell[uu]] += inject_sum(io, s, uu, io.j);
}

// The sums k_inject adds, for the recorded levels of a blocked reverse launch:
// out[(k * n_shots + s) * inj_max + u] for level k = blockIdx.z (j[k] < 0: not recorded).
__global__ void __launch_bounds__(256) k_inject_values(InjectIO<float> io, float* out,
                                                       int n_shots, int inj_max,
                                                       LevelRecords lj) {
  const int s = blockIdx.y, k = blockIdx.z;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (lj.j[k] < 0 || u >= io.inj_count[s]) return;
  out[((int64_t)k * n_shots + s) * inj_max + u] = inject_sum(io, s, io.inj_first[s] + u, lj.j[k]);
}

// ------------------------------------------------------------------------------------------
// reverse step: lam^n from lam^{n+1}, lam^{n+2} (transpose of k_step), plus the imaging sum
// ------------------------------------------------------------------------------------------
template <typename T, int R, int ABS>
__global__ void __launch_bounds__(256) k_adjoint(Geo<T> g, AdjIO<T> io) {
  int64_t s, co[3], idx;
  if (!cell_of(g, s, co, idx)) return;
  const int64_t base = s * g.cells;
  const T* __restrict__ L1 = io.lam1 + base;
  const T* __restrict__ MA = io.ma;
  T a_eff, b, tap;
  damping<T, ABS>(g, co, a_eff, b, tap);
  const T rbar = a_eff * L1[idx];
  // adjoint of the p^{n-1} slot of step n+1: -b * t_p(n+1) * a_eff * lam^{n+2}
  const T tp1 = (ABS == NW_CERJAN && io.taper_next) ? tap : T(1);
  const T pbar_next = -(b * tp1) * (a_eff * io.lam2[base + idx]);
  auto q = [&](int64_t j) { return MA[j] * L1[j]; };
  T zz = T(0);
  if (ABS == NW_PML && g.nd == 2) zz = g.dt2 * g.zc[1][co[1]] * g.zc[2][co[2]];
  T lam = pbar_next + (T(2) - zz) * rbar + lap_of<T, R>(g, co, idx, q);

  T lap_p = T(0);
  const T* __restrict__ Pn = io.p_n ? io.p_n + base : nullptr;
  if (Pn) lap_p = lap_of<T, R>(g, co, idx, [&](int64_t j) { return Pn[j]; });

  if (ABS == NW_PML) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!g.used[a]) continue;
      const int64_t st = stride_of(g, a), i = co[a], n = g.D[a];
      const T qc = q(idx);
      const T qm = i >= 1 ? q(idx - st) : T(0);
      const T qp = i + 1 < n ? q(idx + st) : T(0);
      const T dq0 = qc - qm, dq1 = qp - qc;
      T other = T(0);
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
        if (bb != a && g.used[bb]) other += g.zc[bb][co[bb]];
      const int64_t f0 = face_at(g, a, co, i), f1 = face_at(g, a, co, i + 1);
      const int64_t fb = s * g.fcells[a];
      const T t0 = io.psib_in[a][fb + f0] - g.inv2h[a] * dq0;
      const T t1 = io.psib_in[a][fb + f1] - g.inv2h[a] * dq1;
      const T b0 = (other - g.zf[a][i]) * g.kappa[a][i];
      const T b1 = (other - g.zf[a][i + 1]) * g.kappa[a][i + 1];
      lam -= b1 * t1 - b0 * t0;
      io.psib_out[a][fb + f0] = g.alpha[a][i] * t0 - g.inv2h[a] * dq0;
      if (i == n - 1) io.psib_out[a][fb + f1] = g.alpha[a][i + 1] * t1 - g.inv2h[a] * dq1;
      if (Pn) {
        const T* ph = io.psi_half[a] + fb;  // psi^{n+1/2}
        const T* pl = io.psi_low[a] + fb;   // psi^{n-1/2}
        lap_p += ((ph[f1] + pl[f1]) - (ph[f0] + pl[f0])) * g.inv2h[a];
      }
    }
  }
  io.lam0[base + idx] = lam;
  if (Pn && io.grad) io.grad[base + idx] += (double)rbar * (double)lap_p;
}

// ------------------------------------------------------------------------------------------
// model, gradient fold-back, initial-value state, utilities
// ------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int64_t clamp_field(int64_t i, int64_t a, int64_t n) {
  const int64_t j = i - a;
  return j < 0 ? 0 : (j >= n ? n - 1 : j);
}

template <typename T, int ABS>
__global__ void __launch_bounds__(256) k_model(Geo<T> g, const T* __restrict__ c, T* m, T* ma) {
  int64_t s, co[3], idx;
  if (!cell_of(g, s, co, idx)) return;
  const int64_t i0 = clamp_field<T>(co[0], g.nabs[0], g.F[0]);
  const int64_t i1 = clamp_field<T>(co[1], g.nabs[1], g.F[1]);
  const int64_t i2 = clamp_field<T>(co[2], g.nabs[2], g.F[2]);
  const T cv = c[(i0 * g.F[1] + i1) * g.F[2] + i2];
  const T mv = (cv * cv) * g.dt2;
  m[idx] = mv;
  if (ma) {
    T a_eff, b, tap;
    damping<T, ABS>(g, co, a_eff, b, tap);
    ma[idx] = mv * a_eff;
  }
}

// range of padded cells along one axis that replicate-padding maps onto field cell i
__device__ __forceinline__ void fold_range(int64_t i, int64_t a, int64_t n, int64_t& lo,
                                           int64_t& hi) {
  lo = i == 0 ? 0 : a + i;
  hi = i == n - 1 ? 2 * a + n - 1 : a + i;
}

template <typename T>
__global__ void __launch_bounds__(256) k_model_grad(Geo<T> g, const T* __restrict__ c,
                                                    const double* __restrict__ gm, T* gc,
                                                    int64_t shots) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nf = g.F[0] * g.F[1] * g.F[2];
  if (t >= nf) return;
  const int64_t i2 = t % g.F[2], i1 = (t / g.F[2]) % g.F[1], i0 = t / (g.F[1] * g.F[2]);
  int64_t lo0, hi0, lo1, hi1, lo2, hi2;
  fold_range(i0, g.nabs[0], g.F[0], lo0, hi0);
  fold_range(i1, g.nabs[1], g.F[1], lo1, hi1);
  fold_range(i2, g.nabs[2], g.F[2], lo2, hi2);
  double acc = 0.0;
  for (int64_t s = 0; s < shots; ++s) {
    const double* G = gm + s * g.cells;
    for (int64_t z = lo0; z <= hi0; ++z)
      for (int64_t y = lo1; y <= hi1; ++y)
        for (int64_t x = lo2; x <= hi2; ++x) acc += G[(z * g.D[1] + y) * g.P2 + x];
  }
  gc[t] = (T)(2.0 * (double)g.dt2_d * (double)c[t] * acc);
}

template <typename T>
__device__ __forceinline__ T p0_at(const Geo<T>& g, const T* p0s, int64_t j) {
  const int64_t x = j % g.P2, y = (j / g.P2) % g.D[1], z = j / (g.D[1] * g.P2);
  const int64_t i0 = z - g.nabs[0], i1 = y - g.nabs[1], i2 = x - g.nabs[2];
  if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= g.F[0] || i1 >= g.F[1] || i2 >= g.F[2]) return T(0);
  return p0s[(i0 * g.F[1] + i1) * g.F[2] + i2];
}

template <typename T, int R>
__global__ void __launch_bounds__(256) k_ic_state(Geo<T> g, const T* __restrict__ m,
                                                  const T* __restrict__ p0, T* p_prev, T* p) {
  int64_t s, co[3], idx;
  if (!cell_of(g, s, co, idx)) return;
  const int64_t nf = g.F[0] * g.F[1] * g.F[2];
  const T* p0s = p0 + s * nf;
  const T pc = p0_at(g, p0s, idx);
  const T lap = lap_of<T, R>(g, co, idx, [&](int64_t j) { return p0_at(g, p0s, j); });
  p[s * g.cells + idx] = pc;
  p_prev[s * g.cells + idx] = pc + (T(0.5) * m[idx]) * lap;
}

template <typename T, int R>
__global__ void __launch_bounds__(256) k_ic_grad(Geo<T> g, const T* __restrict__ m,
                                                 const T* __restrict__ lam0,
                                                 const T* __restrict__ pbar, T* gp0) {
  int64_t s, co[3], idx;
  if (!cell_of(g, s, co, idx)) return;
  const int64_t i0 = co[0] - g.nabs[0], i1 = co[1] - g.nabs[1], i2 = co[2] - g.nabs[2];
  if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= g.F[0] || i1 >= g.F[1] || i2 >= g.F[2]) return;
  const T* PB = pbar + s * g.cells;
  const T lap = lap_of<T, R>(g, co, idx, [&](int64_t j) { return (T(0.5) * m[j]) * PB[j]; });
  const int64_t nf = g.F[0] * g.F[1] * g.F[2];
  gp0[s * nf + (i0 * g.F[1] + i1) * g.F[2] + i2] = lam0[s * g.cells + idx] + PB[idx] + lap;
}

template <typename T, int ABS>
__global__ void __launch_bounds__(256) k_prev_adjoint(Geo<T> g, const T* __restrict__ lam1,
                                                      T* pbar) {
  int64_t s, co[3], idx;
  if (!cell_of(g, s, co, idx)) return;
  T a_eff, b, tap;
  damping<T, ABS>(g, co, a_eff, b, tap);
  pbar[s * g.cells + idx] = -b * (a_eff * lam1[s * g.cells + idx]);
}

template <typename T, int R>
__global__ void __launch_bounds__(256) k_laplacian(Geo<T> g, const T* __restrict__ p, T* out) {
  int64_t s, co[3], idx;
  if (!cell_of(g, s, co, idx)) return;
  const T* P = p + s * g.cells;
  out[s * g.cells + idx] = lap_of<T, R>(g, co, idx, [&](int64_t j) { return P[j]; });
}

template <typename T>
__global__ void __launch_bounds__(256) k_clamp(int64_t cells, T* p, const int32_t* __restrict__ idx,
                                               const T* __restrict__ val, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  p[(int64_t)blockIdx.y * cells + idx[k]] = val[k];
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
// Neutral-damping interval of axes 1 and 2 of a float plan with a sponge or Cerjan layer, from
// the device profiles: the run of neutral cells around the middle of each axis.
// PML: a cell is neutral along axis a when its damping and cell profile vanish and so do the
// face profiles of both its faces. In a tile of cells neutral along both axes every face-field
// coefficient (zeta_other - zf) kappa is zero, so the face fields there stay at their initial
// zero and enter the neighbours' updates only through zero coefficients.
static int quiet_box(Geo<float>& g) {
  const bool pml = g.damp_kind == NW_PML;
  auto fetch = [](const float* src, int n, std::vector<float>& v) {
    v.resize(n);
    return hipMemcpy(v.data(), src, n * sizeof(float), hipMemcpyDeviceToHost);
  };
  for (int a = 1; a < 3; ++a) {
    const int n = (int)g.D[a];
    std::vector<float> v, zc, zf;
    hipError_t e = fetch(g.damp[a], n, v);
    if (e == hipSuccess && pml) e = fetch(g.zc[a], n, zc);
    if (e == hipSuccess && pml) e = fetch(g.zf[a], n + 1, zf);
    if (e != hipSuccess)
      return fail(NW_ERR_HIP, std::string("reading the damping profile: ") + hipGetErrorString(e));
    const bool sponge = g.damp_kind == NW_SPONGE;
    auto neutral = [&](int i) {
      if (pml) return v[i] == 0.f && zc[i] == 0.f && zf[i] == 0.f && zf[i + 1] == 0.f;
      return sponge ? v[i] == 0.f : v[i] >= 1.f;
    };
    int lo = n / 2, hi = n / 2;
    if (neutral(n / 2)) {
      while (lo > 0 && neutral(lo - 1)) --lo;
      while (hi < n && neutral(hi)) ++hi;
    }
    g.quiet[a][0] = lo;
    g.quiet[a][1] = hi;
  }
  return NW_OK;
}

template <typename T>
Geo<T> make_geo(const nw_desc& d) {
  Geo<T> g{};
  g.nd = d.ndim;
  for (int a = 0; a < 3; ++a) {
    g.D[a] = d.dims[a];
    g.F[a] = d.field_dims[a];
    g.nabs[a] = d.n_abs[a];
    g.used[a] = a >= 3 - d.ndim;
  }
  g.P2 = d.pitch > 0 ? d.pitch : d.dims[2];
  g.cells = g.D[0] * g.D[1] * g.P2;
  g.damp_kind = d.absorbing;
  for (int a = 0; a < 3; ++a) {
    g.quiet[a][0] = 0;
    g.quiet[a][1] = (int)g.D[a];
  }
  g.FP2 = (std::max<int64_t>(g.P2, g.D[2] + 1) + 3) / 4 * 4;
  for (int a = 0; a < 3; ++a) {
    int64_t f = 1;
    for (int b = 0; b < 2; ++b) f *= g.D[b] + (b == a);
    g.fcells[a] = f * (a == 2 ? g.FP2 : g.D[2]);
  }
  // coefficients of nefi/physics/wave.py:121 per axis, summed into the centre in f64
  const double c0 = d.order == 4 ? -2.5 : -2.0;
  const double c1 = d.order == 4 ? 4.0 / 3.0 : 1.0;
  const double c2 = d.order == 4 ? -1.0 / 12.0 : 0.0;
  double cc = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (!g.used[a]) continue;
    const double ih2 = 1.0 / (d.h[a] * d.h[a]);
    cc += c0 * ih2;
    g.c1[a] = (T)(c1 * ih2);
    g.c2[a] = (T)(c2 * ih2);
    g.inv2h[a] = (T)(0.5 / d.h[a]);
  }
  g.cc = (T)cc;
  g.dt2 = (T)(d.dt * d.dt);
  g.dt2_d = d.dt * d.dt;
  for (int a = 0; a < 3; ++a) {
    g.damp[a] = (const T*)d.damp[a];
    g.alpha[a] = (const T*)d.pml_alpha[a];
    g.kappa[a] = (const T*)d.pml_kappa[a];
    g.zf[a] = (const T*)d.pml_zf[a];
    g.zc[a] = (const T*)d.pml_zc[a];
  }
  return g;
}

}  // namespace nw

using namespace nw;

struct nw_plan {
  nw_desc d;
  Geo<float> gf;
  Geo<double> gd;
  // tiles of the fast 2-D kernels: rows per wavefront strip (0 = generic kernels) and shots
  // advanced by one workgroup; [0] forward step, [1] reverse step
  int rows[2];
  int spw[2];
  int xcd_map;  // NW_OPT_XCD_MAP
  int rows_per_iter;  // NW_OPT_ROWS_PER_ITER (forward step)
  int lds_tiles[2];   // NW_OPT_LDS_TILES_STEP / _ADJOINT
  // temporal blocking of the fast forward step (NW_OPT_TIME_BLOCK): depth, the two extra
  // wavefields a blocked sweep alternates with the caller's pair (its halo reads forbid
  // writing in place), and the receiver-tap values of intermediate levels
  int tblock;
  float* tb_state[2];
  float* tapbuf;
  int2* tap_rows;  // per shot: first and last grid row of its receiver taps (blocked sweeps)
  int2* inj_rows;  // per shot: first and last grid row of its injection cells (blocked reverse)
  int keep_lap_opt;  // NW_OPT_KEEP_LAP (-1: the default of the absorbing layer)
  int fast3d;        // NW_OPT_FAST_3D: the 2.5-D tiled 3-D float32 kernels (when they cover it)
  int fast64;        // NW_OPT_FAST_F64: the LDS-tiled 2-D float64 forward step
  // temporal blocking of the fast reverse step (NW_OPT_TIME_BLOCK_ADJOINT): depth, its own two
  // extra wavefields (a reverse sweep may run beside a blocked recompute, which alternates
  // with tb_state), and the residual-injection sums of a block's recorded levels
  int tblock_adj;
  float* tb_adj[2];
  float* injv;
  // work queue of the blocked forward launches (NW_OPT_WORK_QUEUE): items for the current
  // forward tiles and depth (built when they are set), counters, and the device's CU count
  int work_queue;
  int n_cu;
  int4* wq_items;
  int64_t wq_cap;  // items the buffer holds
  int wq_key[3];   // rows, depth, shots per workgroup of the items (0: none built)
  nw::BlockQueue wq;
  // hipGraph capture of whole nw_forward / nw_adjoint calls (NW_OPT_GRAPH)
  int graph_mode;
  hipStream_t capture_stream;
  std::vector<nw::GraphEntry> graphs;
  int64_t graph_serial;        // launch clock of the graph cache (least recently used goes)
  int64_t shot_bytes;          // one wavefield (all shots)
  int64_t face_bytes[3];       // one face field per axis (all shots), 0 for unused axes
};

namespace nw {

static bool recorded(const nw_desc& d, int n) {
  return n % d.substeps == 0 && n / d.substeps < d.n_t;
}

static dim3 block_for(const int64_t* D) { return D[1] > 1 ? dim3(64, 4, 1) : dim3(256, 1, 1); }

static dim3 grid_for(const int64_t* D, int64_t shots, dim3 blk) {
  return dim3((unsigned)((D[2] + blk.x - 1) / blk.x), (unsigned)((D[1] + blk.y - 1) / blk.y),
              (unsigned)(D[0] * shots));
}

static int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(NW_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return NW_OK;
}

#define NW_TRY(x)            \
  do {                       \
    const int rc_ = (x);     \
    if (rc_ != NW_OK) return rc_; \
  } while (0)

#define NW_HIP(call)                                                                   \
  do {                                                                                 \
    const hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) return fail(NW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
static const Geo<T>& geo(const nw_plan* p);
template <>
const Geo<float>& geo<float>(const nw_plan* p) { return p->gf; }
template <>
const Geo<double>& geo<double>(const nw_plan* p) { return p->gd; }

template <typename T>
static SampleIO<T> sample_io(const nw_plan* P, const T* p, T* traces, int j) {
  const nw_desc& d = P->d;
  SampleIO<T> io{};
  io.p = p;
  io.traces = traces;
  io.j = j;
  io.n_t = d.n_t;
  io.taps = d.rec_taps;
  io.rec_first = d.rec_first;
  io.rec_count = d.rec_count;
  io.trace_first = d.trace_first;
  io.rec_idx = d.rec_idx;
  io.rec_w = (const T*)d.rec_w;
  return io;
}

template <typename T>
static int launch_sample(nw_plan* P, const T* p, T* traces, int j, hipStream_t st) {
  const nw_desc& d = P->d;
  if (d.rec_max <= 0) return NW_OK;
  const SampleIO<T> io = sample_io<T>(P, p, traces, j);
  const int threads = d.rec_max * d.rec_taps;
  dim3 grid((threads + 255) / 256, d.n_shots);
  const int pid = prof_begin(NW_K_SAMPLE, st);
  hipLaunchKernelGGL(k_sample<T>, grid, dim3(256), 0, st, geo<T>(P), io);
  prof_end(pid, st);
  return check_launch("k_sample");
}

template <typename T>
static InjectIO<T> inject_io(const nw_plan* P, T* lam, const T* residual, int j) {
  const nw_desc& d = P->d;
  InjectIO<T> io{};
  io.lam = lam;
  io.residual = residual;
  io.j = j;
  io.n_t = d.n_t;
  io.inj_first = d.inj_first;
  io.inj_count = d.inj_count;
  io.inj_cell = d.inj_cell;
  io.inj_off = d.inj_off;
  io.inj_rec = d.inj_rec;
  io.inj_w = (const T*)d.inj_w;
  io.trace_first = d.trace_first;
  return io;
}

template <typename T>
static int launch_inject(nw_plan* P, T* lam, const T* residual, int j, hipStream_t st) {
  const nw_desc& d = P->d;
  if (d.inj_max <= 0) return NW_OK;
  const InjectIO<T> io = inject_io<T>(P, lam, residual, j);
  dim3 grid((d.inj_max + 255) / 256, d.n_shots);
  const int pid = prof_begin(NW_K_INJECT, st);
  hipLaunchKernelGGL(k_inject<T>, grid, dim3(256), 0, st, geo<T>(P), io);
  prof_end(pid, st);
  return check_launch("k_inject");
}

template <typename T, int R, int ABS>
static void launch_step(nw_plan* P, const StepIO<T>& io, hipStream_t st) {
  const Geo<T>& g = geo<T>(P);
  const dim3 blk = block_for(g.D);
  hipLaunchKernelGGL((k_step<T, R, ABS>), grid_for(g.D, P->d.n_shots, blk), blk, 0, st, g, io);
}

template <typename T, int R, int ABS>
static void launch_adj(nw_plan* P, const AdjIO<T>& io, hipStream_t st) {
  const Geo<T>& g = geo<T>(P);
  const dim3 blk = block_for(g.D);
  hipLaunchKernelGGL((k_adjoint<T, R, ABS>), grid_for(g.D, P->d.n_shots, blk), blk, 0, st, g,
                     io);
}

// whether the forward step of a generic-loop call gathers the receivers of its own input p^n
// (the tiled 3-D step does, in workgroups after its tiles)
template <typename T>
static bool step_samples(const nw_plan* P) {
  return std::is_same<T, float>::value && P->rows[0] == 0 && P->fast3d;
}

template <typename T>
static void dispatch_step(nw_plan* P, const StepIO<T>& io, hipStream_t st,
                          const SampleIO<T>* smp = nullptr) {
  if constexpr (std::is_same<T, float>::value) {
    if (P->rows[0] > 0) {
      fast2d_step(P->gf, P->d.order, P->d.absorbing, io, P->d.n_shots, P->rows[0], P->spw[0],
                  P->xcd_map | (P->rows_per_iter == 2 ? 2 : 0) | (P->lds_tiles[0] ? 4 : 0), st);
      return;
    }
    if (P->fast3d) {
      fast3d_step(P->gf, P->d.order, P->d.absorbing, io, P->d.n_shots, P->xcd_map, smp,
                  P->d.rec_max, st);
      return;
    }
  } else {
    if (P->fast64 && !io.lap_out) {
      fast2d64_step(P->gd, P->d.order, P->d.absorbing, io, P->d.n_shots, P->xcd_map, st);
      return;
    }
  }
  const int r = P->d.order == 4 ? 2 : 1;
  switch (P->d.absorbing * 2 + (r - 1)) {
    case NW_NONE * 2 + 0: launch_step<T, 1, NW_NONE>(P, io, st); break;
    case NW_NONE * 2 + 1: launch_step<T, 2, NW_NONE>(P, io, st); break;
    case NW_SPONGE * 2 + 0: launch_step<T, 1, NW_SPONGE>(P, io, st); break;
    case NW_SPONGE * 2 + 1: launch_step<T, 2, NW_SPONGE>(P, io, st); break;
    case NW_CERJAN * 2 + 0: launch_step<T, 1, NW_CERJAN>(P, io, st); break;
    case NW_CERJAN * 2 + 1: launch_step<T, 2, NW_CERJAN>(P, io, st); break;
    case NW_PML * 2 + 0: launch_step<T, 1, NW_PML>(P, io, st); break;
    default: launch_step<T, 2, NW_PML>(P, io, st); break;
  }
}

template <typename T>
static void dispatch_adj(nw_plan* P, const AdjIO<T>& io, hipStream_t st) {
  if constexpr (std::is_same<T, float>::value) {
    if (P->rows[1] > 0) {
      fast2d_adjoint(P->gf, P->d.order, P->d.absorbing, io, P->d.n_shots, P->rows[1], P->spw[1],
                     P->xcd_map | (P->lds_tiles[1] ? 4 : 0), st);
      return;
    }
    if (P->fast3d && !io.p_n) {  // no imaging term, or one from kept Laplacians
      fast3d_adjoint(P->gf, P->d.order, P->d.absorbing, io, P->d.n_shots, P->xcd_map, st);
      return;
    }
  }
  const int r = P->d.order == 4 ? 2 : 1;
  switch (P->d.absorbing * 2 + (r - 1)) {
    case NW_NONE * 2 + 0: launch_adj<T, 1, NW_NONE>(P, io, st); break;
    case NW_NONE * 2 + 1: launch_adj<T, 2, NW_NONE>(P, io, st); break;
    case NW_SPONGE * 2 + 0: launch_adj<T, 1, NW_SPONGE>(P, io, st); break;
    case NW_SPONGE * 2 + 1: launch_adj<T, 2, NW_SPONGE>(P, io, st); break;
    case NW_CERJAN * 2 + 0: launch_adj<T, 1, NW_CERJAN>(P, io, st); break;
    case NW_CERJAN * 2 + 1: launch_adj<T, 2, NW_CERJAN>(P, io, st); break;
    case NW_PML * 2 + 0: launch_adj<T, 1, NW_PML>(P, io, st); break;
    default: launch_adj<T, 2, NW_PML>(P, io, st); break;
  }
}

// pointers of the PML face fields inside a keep slot
template <typename T>
static void keep_psi(const nw_plan* P, T* slot, T** psi) {
  char* q = (char*)slot + P->shot_bytes;
  for (int a = 0; a < 3; ++a) {
    psi[a] = nullptr;
    if (P->face_bytes[a]) {
      psi[a] = (T*)q;
      q += P->face_bytes[a];
    }
  }
}

// Leave (prev, cur) of a finished sweep in the caller's pair (s0, s1): a pointer swap when
// both already live there, copies for the ones that live in keep slots or plan buffers.
template <typename T>
static int hand_back(nw_plan* P, void** state, T* s0, T* s1, T* prev, T* cur, hipStream_t st) {
  const bool prev_own = prev == s0 || prev == s1;
  const bool cur_own = cur == s0 || cur == s1;
  if (prev_own && cur_own) {
    state[0] = prev;
    state[1] = cur;
  } else if (cur_own) {
    T* other = cur == s0 ? s1 : s0;
    NW_HIP(hipMemcpyAsync(other, prev, P->shot_bytes, hipMemcpyDeviceToDevice, st));
    state[0] = other;
    state[1] = cur;
  } else if (prev_own) {
    T* other = prev == s0 ? s1 : s0;
    NW_HIP(hipMemcpyAsync(other, cur, P->shot_bytes, hipMemcpyDeviceToDevice, st));
    state[0] = prev;
    state[1] = other;
  } else {
    NW_HIP(hipMemcpyAsync(s0, prev, P->shot_bytes, hipMemcpyDeviceToDevice, st));
    NW_HIP(hipMemcpyAsync(s1, cur, P->shot_bytes, hipMemcpyDeviceToDevice, st));
    state[0] = s0;
    state[1] = s1;
  }
  return NW_OK;
}

// Keep buffers hold the Laplacian of p^n (instead of p^n) when both fast 2-D kernel families
// run: the forward kernels compute it anyway, and the reverse kernels' imaging term then reads
// it pointwise instead of re-applying the stencil to a p tile (PML: and instead of reading both
// levels of both face fields, so there it is the default).
static bool keep_lap(const nw_plan* P) {
  const bool tiled2d = P->rows[0] > 0 && P->rows[1] > 0;
  const bool tiled3d = P->fast3d != 0;
  const bool want = P->keep_lap_opt < 0 ? (P->d.absorbing == NW_PML || tiled3d)
                                        : P->keep_lap_opt != 0;
  return P->d.dtype == NW_F32 && want && (tiled2d || tiled3d);
}

// The blocked sweep applies when the plan asks for it and the forward tiles have a blocked
// kernel; a sweep recording traces also needs every shot's receiver taps in one scan.
static bool use_blocked(const nw_plan* P, bool traces) {
  const nw_desc& d = P->d;
  return d.dtype == NW_F32 && d.absorbing != NW_PML && P->tblock > 1 && P->rows[0] > 0 &&
         P->tb_state[0] &&
         fast2d_block_supported(P->rows[0], P->tblock) &&
         (!traces || d.rec_max * d.rec_taps <= 256 * kTapChunks);
}

// Forward sweep in blocks of up to tblock steps (2-D float32 fast path). Blocks start at every
// checkpoint step, so the kernel writes a checkpoint from the block's inputs. Outside keep
// mode the block's last two levels go to the two of {s0, s1, tb_state} the block does not read.
static int forward_blocked(nw_plan* P, const float* m, void** state, int n_begin, int n_end,
                           float* traces, float* ckpt, int ckpt_every, float* keep,
                           int keep_begin, int first_step, hipStream_t st) {
  const nw_desc& d = P->d;
  const int64_t keep_slot = nw_keep_slot_bytes(P);
  const int64_t ck_slot = nw_ckpt_slot_bytes(P);
  float* s0 = (float*)state[0];
  float* s1 = (float*)state[1];
  float* const bufs[4] = {s0, s1, P->tb_state[0], P->tb_state[1]};
  auto kslot = [&](int n) {
    return (float*)((char*)keep + (int64_t)(n - keep_begin) * keep_slot);
  };
  const bool keeping = keep != nullptr;
  const bool klap = keeping && keep_lap(P);
  if (keeping && !klap && keep_begin == n_begin)
    NW_HIP(hipMemcpyAsync(kslot(n_begin), s1, P->shot_bytes, hipMemcpyDeviceToDevice, st));
  const bool taps = traces && d.rec_max > 0;
  float* prev = s0;
  float* cur = s1;
  for (int n = n_begin; n < n_end;) {
    if (traces && recorded(d, n)) NW_TRY(launch_sample<float>(P, cur, traces, n / d.substeps, st));
    int L = std::min(P->tblock, n_end - n);
    if (ckpt && ckpt_every > 0) L = std::min(L, (n / ckpt_every + 1) * ckpt_every - n);
    BlockIO io{};
    io.m = m;
    io.p_prev = prev;
    io.p = cur;
    io.levels = L;
    io.n0 = n;
    io.first_step = first_step;
    io.n_src = d.n_steps;
    if (d.src_taps > 0) {
      io.src_taps = d.src_taps;
      io.src_idx = d.src_idx;
      io.src_val = (const float*)d.src_val;
      io.wavelet = (const float*)d.wavelet;
    }
    if (ckpt && ckpt_every > 0 && n % ckpt_every == 0) {
      float* slot = (float*)((char*)ckpt + (int64_t)(n / ckpt_every) * ck_slot);
      io.ck_prev = slot;
      io.ck_cur = (float*)((char*)slot + P->shot_bytes);
    }
    if (klap)
      for (int k = 1; k <= L; ++k) io.lap_out[k] = kslot(n + k - 1);
    if (keeping && !klap) {
      for (int k = 1; k <= L; ++k) io.out[k] = kslot(n + k);
    } else {
      float* spare[2] = {nullptr, nullptr};
      int ns = 0;
      for (float* b : bufs)
        if (b != prev && b != cur && ns < 2) spare[ns++] = b;
      if (L >= 2) {
        io.out[L - 1] = spare[0];
        io.out[L] = spare[1];
      } else {
        io.out[1] = spare[0];
      }
    }
    if (taps && L >= 2) {
      io.tapbuf = P->tapbuf;
      io.tap_stride = d.rec_max * d.rec_taps;
      io.taps = d.rec_taps;
      io.rec_first = d.rec_first;
      io.rec_count = d.rec_count;
      io.rec_idx = d.rec_idx;
      io.tap_rows = reinterpret_cast<const int32_t*>(P->tap_rows);
    }
    const int pid = prof_begin(NW_K_STEPS, st, L);
    // keep sweeps run beside reverse steps in an overlapped reverse sweep, where persistent
    // workgroups would hold their CUs until the queue is empty: by default they take one
    // workgroup per tile
    const bool queued = (P->work_queue == 2 || (P->work_queue == 1 && !keeping)) &&
                        P->wq_key[0] == P->rows[0] && P->wq_key[1] == P->tblock &&
                        P->wq_key[2] == P->spw[0];
    fast2d_steps(P->gf, d.order, d.absorbing, io, d.n_shots, P->rows[0], P->tblock, P->spw[0],
                 P->xcd_map, queued ? &P->wq : nullptr, st);
    prof_end(pid, st);
    NW_TRY(check_launch("k4_steps"));
    if (io.tapbuf) {
      LevelRecords lj;
      int nrec = 0;
      for (int k = 0; k < kMaxDepth; ++k) lj.j[k] = -1;
      for (int k = 1; k < L; ++k)
        if (recorded(d, n + k)) {
          lj.j[k - 1] = (n + k) / d.substeps;
          ++nrec;
        }
      if (nrec > 0) {
        const int pid2 = prof_begin(NW_K_SAMPLE, st, nrec);
        fast2d_sample_levels(P->gf, sample_io<float>(P, nullptr, traces, 0), P->tapbuf,
                             io.tap_stride, d.rec_max, d.n_shots, L, lj, st);
        prof_end(pid2, st);
        NW_TRY(check_launch("k_sample_levels"));
      }
    }
    prev = L >= 2 ? io.out[L - 1] : cur;
    cur = io.out[L];
    n += L;
  }
  if (traces && recorded(d, n_end)) NW_TRY(launch_sample<float>(P, cur, traces, n_end / d.substeps, st));
  if (n_end > n_begin) NW_TRY(hand_back<float>(P, state, s0, s1, prev, cur, st));
  return NW_OK;
}

template <typename T>
static int forward_impl(nw_plan* P, const T* m, void** state, void** psi, int n_begin, int n_end,
                        T* traces, T* ckpt, int ckpt_every, T* keep, int keep_begin,
                        int first_step, hipStream_t st) {
  const nw_desc& d = P->d;
  if constexpr (std::is_same<T, float>::value) {
    if (use_blocked(P, traces != nullptr))
      return forward_blocked(P, m, state, n_begin, n_end, traces, ckpt, ckpt_every, keep,
                             keep_begin, first_step, st);
  }
  const bool pml = d.absorbing == NW_PML;
  const int64_t keep_slot = nw_keep_slot_bytes(P);
  const int64_t ck_slot = nw_ckpt_slot_bytes(P);
  T* s0 = (T*)state[0];
  T* s1 = (T*)state[1];
  // psi pairs per 3-D axis: current and spare
  T* pa[3] = {nullptr, nullptr, nullptr};
  T* pb[3] = {nullptr, nullptr, nullptr};
  if (pml) {
    int k = 0;
    for (int a = 0; a < 3; ++a)
      if (P->face_bytes[a]) {
        pa[a] = (T*)psi[k];
        pb[a] = (T*)psi[d.ndim + k];
        ++k;
      }
  }
  auto kslot = [&](int n) { return (T*)((char*)keep + (int64_t)(n - keep_begin) * keep_slot); };
  const bool klap = keep != nullptr && keep_lap(P);
  const bool keeping = keep != nullptr && !klap;  // keep slots carry the state chain
  if (keeping && keep_begin == n_begin) {
    NW_HIP(hipMemcpyAsync(kslot(n_begin), s1, P->shot_bytes, hipMemcpyDeviceToDevice, st));
    if (pml) {
      T* kp[3];
      keep_psi(P, kslot(n_begin), kp);
      for (int a = 0; a < 3; ++a)
        if (pa[a]) NW_HIP(hipMemcpyAsync(kp[a], pa[a], P->face_bytes[a], hipMemcpyDeviceToDevice, st));
    }
  }
  // current pointers (may point into keep slots)
  T* prev = s0;
  T* cur = s1;
  T* psi_cur[3] = {pa[0], pa[1], pa[2]};
  const bool fused_sample = step_samples<T>(P);
  for (int n = n_begin; n < n_end; ++n) {
    const bool sample_n = traces && recorded(d, n);
    if (sample_n && !fused_sample)
      NW_TRY(launch_sample<T>(P, cur, traces, n / d.substeps, st));
    StepIO<T> io{};
    io.m = m;
    io.p_prev = prev;
    io.p = cur;
    T* next;
    T* psi_next[3] = {nullptr, nullptr, nullptr};
    if (keeping) {
      next = kslot(n + 1);
      if (pml) keep_psi(P, next, psi_next);
    } else {
      next = (prev == s0 || prev == s1) ? prev : (cur == s0 ? s1 : s0);
      for (int a = 0; a < 3; ++a)
        if (pa[a]) psi_next[a] = (psi_cur[a] == pa[a]) ? pb[a] : pa[a];
    }
    io.p_next = next;
    for (int a = 0; a < 3; ++a) {
      io.psi_old[a] = psi_cur[a];
      io.psi_new[a] = psi_next[a];
    }
    io.taper_prev = n > first_step;
    if (klap) io.lap_out = kslot(n);
    if (d.src_taps > 0 && n < d.n_steps) {
      io.src_taps = d.src_taps;
      io.src_idx = d.src_idx;
      io.src_val = (const T*)d.src_val;
      io.wavelet = (const T*)d.wavelet;
      io.n = n;
    }
    if (ckpt && ckpt_every > 0 && n % ckpt_every == 0) {
      T* slot = (T*)((char*)ckpt + (int64_t)(n / ckpt_every) * ck_slot);
      io.ck_prev = slot;
      io.ck_cur = (T*)((char*)slot + P->shot_bytes);
      char* q = (char*)slot + 2 * P->shot_bytes;
      for (int a = 0; a < 3; ++a)
        if (P->face_bytes[a]) {
          io.ck_psi[a] = (T*)q;
          q += P->face_bytes[a];
        }
    }
    const int pid = prof_begin(NW_K_STEP, st);
    if (sample_n && fused_sample) {
      const SampleIO<T> smp = sample_io<T>(P, cur, traces, n / d.substeps);
      dispatch_step<T>(P, io, st, &smp);
    } else {
      dispatch_step<T>(P, io, st);
    }
    prof_end(pid, st);
    NW_TRY(check_launch("k_step"));
    prev = cur;
    cur = next;
    for (int a = 0; a < 3; ++a) psi_cur[a] = psi_next[a];
  }
  if (traces && recorded(d, n_end)) NW_TRY(launch_sample<T>(P, cur, traces, n_end / d.substeps, st));
  // hand the final state back in the caller's buffers
  if (n_end > n_begin) {
    NW_TRY(hand_back<T>(P, state, s0, s1, prev, cur, st));
    if (pml) {
      int k = 0;
      for (int a = 0; a < 3; ++a) {
        if (!pa[a]) continue;
        T* c = psi_cur[a];
        if (c != pa[a] && c != pb[a]) {
          NW_HIP(hipMemcpyAsync(pa[a], c, P->face_bytes[a], hipMemcpyDeviceToDevice, st));
          c = pa[a];
        }
        psi[k] = c;
        psi[d.ndim + k] = c == pa[a] ? pb[a] : pa[a];
        ++k;
      }
    }
  }
  return NW_OK;
}

// The blocked reverse sweep applies when the plan asks for it, the reverse tiles have a
// blocked kernel, the imaging term (if any) reads kept Laplacians, and a shot's injection cells
// fit one scan of the workgroup.
static bool use_blocked_adj(const nw_plan* P, const void* keep, const void* residual) {
  const nw_desc& d = P->d;
  return d.dtype == NW_F32 && d.absorbing != NW_PML && P->tblock_adj > 1 && P->rows[1] > 0 &&
         P->tb_adj[0] &&
         P->injv && fast2d_adj_block_supported(P->rows[1], P->tblock_adj) &&
         (!keep || keep_lap(P)) && (!residual || d.inj_max <= 256 * kTapChunks);
}

// Reverse sweep in blocks of up to tblock_adj steps (2-D float32 fast path). Each block's two
// outputs go to the two of {lam pair, tb_adj} the block does not read.
static int adjoint_blocked(nw_plan* P, const float* ma, void** lam, int n_begin, int n_end,
                           const float* keep, int keep_begin, const float* residual,
                           int inject_end, double* grad, int first_step, hipStream_t st) {
  const nw_desc& d = P->d;
  const int64_t keep_slot = nw_keep_slot_bytes(P);
  auto kslot = [&](int n) {
    return (const float*)((const char*)keep + (int64_t)(n - keep_begin) * keep_slot);
  };
  float* a0 = (float*)lam[0];
  float* a1 = (float*)lam[1];
  float* const bufs[4] = {a0, a1, P->tb_adj[0], P->tb_adj[1]};
  float* L1 = a0;  // lambda^{n}
  float* L2 = a1;  // lambda^{n+1}
  if (residual && inject_end && recorded(d, n_end))
    NW_TRY(launch_inject<float>(P, L1, residual, n_end / d.substeps, st));
  for (int n = n_end; n > n_begin;) {
    const int L = std::min(P->tblock_adj, n - n_begin);
    AdjBlockIO io{};
    io.ma = ma;
    io.lam1 = L1;
    io.lam2 = L2;
    io.levels = L;
    io.nh = n;
    io.first_step = first_step;
    if (keep) {
      io.grad = grad;
      for (int k = 1; k <= L; ++k) io.lap[k] = kslot(n - k);
    }
    float* spare[2] = {nullptr, nullptr};
    int ns = 0;
    for (float* b : bufs)
      if (b != L1 && b != L2 && ns < 2) spare[ns++] = b;
    if (L >= 2) {
      io.out[L - 1] = spare[0];
      io.out[L] = spare[1];
    } else {
      io.out[1] = spare[0];
    }
    if (residual && d.inj_max > 0) {
      LevelRecords lj;
      int nrec = 0;
      for (int k = 0; k < kMaxDepth; ++k) lj.j[k] = -1;
      for (int k = 1; k <= L; ++k)
        if (recorded(d, n - k)) {
          lj.j[k - 1] = (n - k) / d.substeps;
          io.inj_rec[k] = 1;
          ++nrec;
        }
      if (nrec > 0) {
        const InjectIO<float> iio = inject_io<float>(P, nullptr, residual, 0);
        const dim3 grid((d.inj_max + 255) / 256, d.n_shots, L);
        const int pid = prof_begin(NW_K_INJECT, st, nrec);
        hipLaunchKernelGGL(k_inject_values, grid, dim3(256), 0, st, iio, P->injv, d.n_shots,
                           d.inj_max, lj);
        prof_end(pid, st);
        NW_TRY(check_launch("k_inject_values"));
        io.injv = P->injv;
        io.inj_max = d.inj_max;
        io.inj_first = d.inj_first;
        io.inj_count = d.inj_count;
        io.inj_cell = d.inj_cell;
        io.inj_rows = reinterpret_cast<const int32_t*>(P->inj_rows);
      }
    }
    const int pid = prof_begin(NW_K_ADJOINT_STEPS, st, L);
    fast2d_adjoint_steps(P->gf, d.order, d.absorbing, io, d.n_shots, P->rows[1], P->tblock_adj,
                         P->spw[1], P->xcd_map, st);
    prof_end(pid, st);
    NW_TRY(check_launch("k5_adjoint"));
    L2 = L >= 2 ? io.out[L - 1] : L1;
    L1 = io.out[L];
    n -= L;
  }
  return hand_back<float>(P, lam, a0, a1, L1, L2, st);
}

template <typename T>
static int adjoint_impl(nw_plan* P, const T* m, const T* ma, void** lam, void** psib, int n_begin,
                        int n_end, const T* keep, int keep_begin, const T* residual,
                        int inject_end, double* grad, int first_step, hipStream_t st) {
  (void)m;
  const nw_desc& d = P->d;
  if constexpr (std::is_same<T, float>::value) {
    if (use_blocked_adj(P, keep, residual))
      return adjoint_blocked(P, ma, lam, n_begin, n_end, keep, keep_begin, residual, inject_end,
                             grad, first_step, st);
  }
  const bool pml = d.absorbing == NW_PML;
  const int64_t keep_slot = nw_keep_slot_bytes(P);
  T* L1 = (T*)lam[0];
  T* L2 = (T*)lam[1];
  T* qa[3] = {nullptr, nullptr, nullptr};
  T* qb[3] = {nullptr, nullptr, nullptr};
  if (pml) {
    int k = 0;
    for (int a = 0; a < 3; ++a)
      if (P->face_bytes[a]) {
        qa[a] = (T*)psib[k];
        qb[a] = (T*)psib[d.ndim + k];
        ++k;
      }
  }
  auto kslot = [&](int n) {
    return (const T*)((const char*)keep + (int64_t)(n - keep_begin) * keep_slot);
  };
  if (residual && inject_end && recorded(d, n_end))
    NW_TRY(launch_inject<T>(P, L1, residual, n_end / d.substeps, st));
  for (int n = n_end - 1; n >= n_begin; --n) {
    AdjIO<T> io{};
    io.ma = ma;
    io.lam1 = L1;
    io.lam2 = L2;
    io.lam0 = L2;
    io.taper_next = (n + 1) > first_step;
    if (keep) {
      if (keep_lap(P))
        io.lap_n = kslot(n);
      else
        io.p_n = kslot(n);
      io.grad = grad;
      if (pml && !keep_lap(P)) {
        T* lo[3];
        T* hi[3];
        keep_psi(P, (T*)kslot(n), lo);
        keep_psi(P, (T*)kslot(n + 1), hi);
        for (int a = 0; a < 3; ++a) {
          io.psi_low[a] = lo[a];
          io.psi_half[a] = hi[a];
        }
      }
    }
    for (int a = 0; a < 3; ++a) {
      io.psib_in[a] = qa[a];
      io.psib_out[a] = qb[a];
    }
    const int pid = prof_begin(NW_K_ADJOINT, st);
    dispatch_adj<T>(P, io, st);
    prof_end(pid, st);
    NW_TRY(check_launch("k_adjoint"));
    T* t = L1;
    L1 = L2;  // now lam^n
    L2 = t;   // lam^{n+1}
    for (int a = 0; a < 3; ++a) {
      T* u = qa[a];
      qa[a] = qb[a];
      qb[a] = u;
    }
    if (residual && recorded(d, n)) NW_TRY(launch_inject<T>(P, L1, residual, n / d.substeps, st));
  }
  lam[0] = L1;
  lam[1] = L2;
  if (pml) {
    int k = 0;
    for (int a = 0; a < 3; ++a)
      if (qa[a]) {
        psib[k] = qa[a];
        psib[d.ndim + k] = qb[a];
        ++k;
      }
  }
  return NW_OK;
}

// plans alive in the process (for timing reads across their captured graphs)
std::vector<nw_plan*> g_plans;

// Graphs a plan keeps. One gradient replays about 2 x (blocks) + 1 distinct calls, so 256
// holds the calls of sweeps up to ~120 snapshot blocks before any is captured twice.
constexpr size_t kMaxGraphs = 256;

// Destroy one entry after its latest launch has completed (a launch may still be queued on the
// caller's stream).
static void destroy_entry(GraphEntry& e) {
  if (e.done) {
    (void)hipEventSynchronize(e.done);  // cleanup: nothing to report
    (void)hipEventDestroy(e.done);
  }
  if (e.exec) (void)hipGraphExecDestroy(e.exec);
  e.done = nullptr;
  e.exec = nullptr;
}

static void drop_graphs(nw_plan* P) {
  for (auto& e : P->graphs) destroy_entry(e);
  P->graphs.clear();
}

// Make room for one more entry by destroying the least recently launched one.
static void evict_lru(nw_plan* P) {
  if (P->graphs.size() < kMaxGraphs) return;
  size_t victim = 0;
  for (size_t i = 1; i < P->graphs.size(); ++i)
    if (P->graphs[i].last_use < P->graphs[victim].last_use) victim = i;
  destroy_entry(P->graphs[victim]);
  P->graphs.erase(P->graphs.begin() + (std::ptrdiff_t)victim);
}

// Run `body(stream)` directly, or (graph mode) capture it once per distinct key on the plan's
// capture stream and replay the instantiated graph on `st`. `io` holds the pointer arrays the
// call rearranges; replays restore the arrangement recorded at capture.
template <typename F>
static int run_captured(nw_plan* P, std::vector<uintptr_t> key, std::vector<void**> io_arrays,
                        std::vector<int> io_sizes, hipStream_t st, F body) {
  // timed calls run eagerly: events recorded inside a graph cannot be read back here
  if (!P->graph_mode || g_prof.on) return body(st);
  for (size_t a = 0; a < io_arrays.size(); ++a)
    for (int i = 0; i < io_sizes[a]; ++i) key.push_back((uintptr_t)io_arrays[a][i]);
  GraphEntry* hit = nullptr;
  for (auto& e : P->graphs)
    if (e.key == key) hit = &e;
  if (!hit) {
    if (!P->capture_stream) NW_HIP(hipStreamCreateWithFlags(&P->capture_stream, hipStreamNonBlocking));
    evict_lru(P);
    NW_HIP(hipStreamBeginCapture(P->capture_stream, hipStreamCaptureModeThreadLocal));
    const int rc = body(P->capture_stream);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(P->capture_stream, &graph);
    if (rc != NW_OK || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc != NW_OK) return rc;
      return fail(NW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    }
    GraphEntry e;
    e.key = key;
    const hipError_t ei = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess)
      return fail(NW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    for (size_t a = 0; a < io_arrays.size(); ++a)
      for (int i = 0; i < io_sizes[a]; ++i) e.out.push_back(io_arrays[a][i]);
    P->graphs.push_back(std::move(e));
    hit = &P->graphs.back();
  }
  NW_HIP(hipGraphLaunch(hit->exec, st));
  if (!hit->done) NW_HIP(hipEventCreateWithFlags(&hit->done, hipEventDisableTiming));
  NW_HIP(hipEventRecord(hit->done, st));
  hit->last_use = ++P->graph_serial;
  size_t k = 0;
  for (size_t a = 0; a < io_arrays.size(); ++a)
    for (int i = 0; i < io_sizes[a]; ++i) io_arrays[a][i] = hit->out[k++];
  return NW_OK;
}

}  // namespace nw

// the tiled kernels move 16-byte vectors: every device buffer they touch is 16-byte aligned
static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

static bool all_aligned16(void* const* bufs, int n) {
  for (int i = 0; i < n; ++i)
    if (bufs[i] && !aligned16(bufs[i])) return false;
  return true;
}

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
extern "C" {

int32_t nw_abi_version(void) { return NW_ABI_VERSION; }

int64_t nw_desc_bytes(void) { return (int64_t)sizeof(nw_desc); }

int32_t nw_profile_enable(int32_t on) {
  g_prof.recycle();
  g_prof.on = on != 0;
  return NW_OK;
}

int32_t nw_profile_read(int32_t kind, double* total_ms, int64_t* launches) {
  if (!total_ms || !launches) return fail(NW_ERR_ARG, "nw_profile_read: null argument");
  double ms = 0.0;
  int64_t n = 0;
  auto add = [&](const ProfRec& r, int64_t times) -> int {
    NW_HIP(hipEventSynchronize(r.stop));
    float t = 0.f;
    NW_HIP(hipEventElapsedTime(&t, r.start, r.stop));
    ms += (double)t * (double)times;
    n += times;
    return NW_OK;
  };
  for (auto& r : g_prof.recs)
    if (r.kind == kind) NW_TRY(add(r, 1));
  *total_ms = ms;
  *launches = n;
  return NW_OK;
}

int32_t nw_profile_units(int32_t kind, int64_t* units) {
  if (!units) return fail(NW_ERR_ARG, "nw_profile_units: null argument");
  int64_t u = 0;
  for (auto& r : g_prof.recs)
    if (r.kind == kind) u += r.units;
  *units = u;
  return NW_OK;
}

const char* nw_last_error(void) { return nw::g_err.c_str(); }

int32_t nw_plan_create(const nw_desc* desc, nw_plan** out) {
  if (!desc || !out) return fail(NW_ERR_ARG, "nw_plan_create: null argument");
  const nw_desc& d = *desc;
  if (d.ndim < 1 || d.ndim > 3) return fail(NW_ERR_ARG, "ndim must be 1..3");
  if (d.order != 2 && d.order != 4) return fail(NW_ERR_ARG, "order must be 2 or 4");
  if (d.absorbing < NW_NONE || d.absorbing > NW_PML) return fail(NW_ERR_ARG, "bad absorbing");
  if (d.dtype != NW_F32 && d.dtype != NW_F64) return fail(NW_ERR_ARG, "bad dtype");
  if (d.absorbing == NW_PML && d.ndim > 2) return fail(NW_ERR_ARG, "PML supports 1-D/2-D");
  if (d.n_shots < 1 || d.substeps < 1 || d.n_t < 1 || d.n_steps < 0)
    return fail(NW_ERR_ARG, "bad time axis or shot count");
  for (int a = 0; a < 3; ++a) {
    if (d.dims[a] < 1) return fail(NW_ERR_ARG, "dims must be >= 1");
    if (a < 3 - d.ndim && d.dims[a] != 1) return fail(NW_ERR_ARG, "unused leading dims must be 1");
    if (d.field_dims[a] + 2 * d.n_abs[a] != d.dims[a])
      return fail(NW_ERR_ARG, "dims must equal field_dims + 2*n_abs");
  }
  if (d.pitch != 0 && d.pitch < d.dims[2]) return fail(NW_ERR_ARG, "pitch must be >= dims[2]");
  const int64_t cells = d.dims[0] * d.dims[1] * (d.pitch > 0 ? d.pitch : d.dims[2]);
  if (cells >= (int64_t)1 << 31) return fail(NW_ERR_ARG, "padded grid exceeds 2^31 cells");
  if (d.dims[0] * d.n_shots > 65535) return fail(NW_ERR_ARG, "dims[0] * n_shots exceeds 65535");
  if (d.rec_max > 0 && (d.rec_taps != (1 << d.ndim) || !d.rec_idx || !d.rec_w))
    return fail(NW_ERR_ARG, "receiver tables inconsistent with ndim");
  if (d.src_taps > 0 && (!d.src_idx || !d.src_val || !d.wavelet))
    return fail(NW_ERR_ARG, "source tables missing");
  for (int a = 0; a < 3; ++a) {
    const bool used = a >= 3 - d.ndim;
    const bool need_damp = used && d.absorbing != NW_NONE;
    if (need_damp && !d.damp[a]) return fail(NW_ERR_ARG, "damping profile missing");
    if (used && d.absorbing == NW_PML &&
        (!d.pml_alpha[a] || !d.pml_kappa[a] || !d.pml_zf[a] || !d.pml_zc[a]))
      return fail(NW_ERR_ARG, "PML profiles missing");
  }
  nw_plan* p = new nw_plan();
  p->d = d;
  p->gf = make_geo<float>(d);
  p->gd = make_geo<double>(d);
  const bool fast = fast2d_supported(p->gf, d.absorbing, d.dtype);
  if (fast && d.absorbing != NW_NONE) {
    const int rc = quiet_box(p->gf);
    if (rc != NW_OK) {
      delete p;
      return rc;
    }
  }
  for (int k = 0; k < 2; ++k) {
    p->rows[k] = fast ? 16 : 0;
    p->spw[k] = 1;
  }
  p->xcd_map = 1;
  p->rows_per_iter = 1;
  p->lds_tiles[0] = p->lds_tiles[1] = 0;
  p->tblock = 1;
  p->tb_state[0] = p->tb_state[1] = nullptr;
  p->tb_adj[0] = p->tb_adj[1] = nullptr;
  p->tapbuf = nullptr;
  p->tap_rows = nullptr;
  p->inj_rows = nullptr;
  p->keep_lap_opt = -1;
  p->fast3d = fast3d_supported(p->gf, d.absorbing, d.dtype) ? 1 : 0;
  p->fast64 = fast2d64_supported(p->gd, d.absorbing, d.dtype) ? 1 : 0;
  p->tblock_adj = 1;
  p->injv = nullptr;
  p->work_queue = 1;
  p->n_cu = 0;
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      p->n_cu = 0;  // no device: blocked launches use one workgroup per tile
  }
  p->wq_items = nullptr;
  p->wq_cap = 0;
  p->wq_key[0] = p->wq_key[1] = p->wq_key[2] = 0;
  p->wq = nw::BlockQueue{};
  p->graph_mode = 0;
  p->capture_stream = nullptr;
  p->graph_serial = 0;
  nw::g_plans.push_back(p);
  const size_t es = d.dtype == NW_F64 ? 8 : 4;
  p->shot_bytes = cells * (int64_t)es * d.n_shots;
  for (int a = 0; a < 3; ++a)
    p->face_bytes[a] = (d.absorbing == NW_PML && a >= 3 - d.ndim)
                           ? (p->gf.fcells[a] * d.n_shots + 63) / 64 * 64 * (int64_t)es
                           : 0;
  *out = p;
  return NW_OK;
}

int32_t nw_plan_destroy(nw_plan* plan) {
  if (!plan) return NW_OK;
  drop_graphs(plan);
  if (plan->capture_stream) (void)hipStreamDestroy(plan->capture_stream);
  for (float* b : {plan->tb_state[0], plan->tb_state[1], plan->tb_adj[0], plan->tb_adj[1],
                   plan->tapbuf, plan->injv})
    if (b) (void)hipFree(b);
  if (plan->tap_rows) (void)hipFree(plan->tap_rows);
  if (plan->inj_rows) (void)hipFree(plan->inj_rows);
  if (plan->wq_items) (void)hipFree(plan->wq_items);
  if (plan->wq.ctr) (void)hipFree(plan->wq.ctr);
  for (size_t i = 0; i < nw::g_plans.size(); ++i)
    if (nw::g_plans[i] == plan) {
      nw::g_plans.erase(nw::g_plans.begin() + i);
      break;
    }
  delete plan;
  return NW_OK;
}

int64_t nw_grad_slices(const nw_plan* P) {
  if (P->d.dtype == NW_F32 && P->rows[1] > 0) return (P->d.n_shots + P->spw[1] - 1) / P->spw[1];
  return P->d.n_shots;
}

// First and last grid row of the cells each shot's table entries [first[s], first[s] +
// count[s]) point at (`per` cells per entry; an empty range for a shot without entries), as a
// device array of one int2 per shot.
static int shot_rows(const nw_plan* P, const int32_t* first_d, const int32_t* count_d,
                     const int32_t* cells_d, int per, int2** out) {
  const int S = P->d.n_shots;
  std::vector<int2> tr(S, make_int2(INT32_MAX, -1));
  if (first_d && count_d && cells_d) {
    std::vector<int32_t> first(S), count(S);
    NW_HIP(hipMemcpy(first.data(), first_d, S * sizeof(int32_t), hipMemcpyDefault));
    NW_HIP(hipMemcpy(count.data(), count_d, S * sizeof(int32_t), hipMemcpyDefault));
    int64_t rows = 0;
    for (int s = 0; s < S; ++s) rows = std::max<int64_t>(rows, (int64_t)first[s] + count[s]);
    std::vector<int32_t> idx((size_t)rows * per);
    if (rows > 0)
      NW_HIP(hipMemcpy(idx.data(), cells_d, idx.size() * sizeof(int32_t), hipMemcpyDefault));
    const int64_t P2 = P->gf.P2;
    for (int s = 0; s < S; ++s)
      for (int64_t e = (int64_t)first[s] * per; e < ((int64_t)first[s] + count[s]) * per; ++e) {
        const int y = (int)(idx[e] / P2);
        tr[s].x = std::min(tr[s].x, y);
        tr[s].y = std::max(tr[s].y, y);
      }
  }
  NW_HIP(hipMalloc((void**)out, S * sizeof(int2)));
  NW_HIP(hipMemcpy(*out, tr.data(), S * sizeof(int2), hipMemcpyHostToDevice));
  return NW_OK;
}

// Row ranges of every shot's receiver taps and injection cells: a blocked forward tile scans a
// shot's taps only when they reach its output rows, a blocked reverse tile its injection cells
// only when they reach its region. Reads the tables back once, when the blocking buffers are
// allocated.
static int build_tap_rows(nw_plan* P) {
  const nw_desc& d = P->d;
  const bool rec = d.rec_max > 0, inj = d.inj_max > 0;
  NW_TRY(shot_rows(P, rec ? d.rec_first : nullptr, rec ? d.rec_count : nullptr,
                   rec ? d.rec_idx : nullptr, d.rec_taps, &P->tap_rows));
  NW_TRY(shot_rows(P, inj ? d.inj_first : nullptr, inj ? d.inj_count : nullptr,
                   inj ? d.inj_cell : nullptr, 1, &P->inj_rows));
  return NW_OK;
}

// (Re)builds the work queue of the blocked forward launches for the plan's current forward
// tiles and depth, or marks it unused (the launches then take one workgroup per tile). Runs
// when tiles or options are set, never inside a captured call: it waits for the device (a
// launch in flight may still read the previous items) and copies the items synchronously.
static int build_queue(nw_plan* P) {
  P->wq_key[0] = P->wq_key[1] = P->wq_key[2] = 0;
  const nw_desc& d = P->d;
  if (!P->work_queue || P->n_cu <= 0 || d.dtype != NW_F32 || d.absorbing == NW_PML ||
      P->tblock < 2 || P->rows[0] <= 0 || !fast2d_block_supported(P->rows[0], P->tblock))
    return NW_OK;
  std::vector<int4> items;
  int off[kQueues + 1], per_cu = 1;
  // the last tiles of every queue go as single shots: about half a queue's CUs' worth
  const int split_tail = std::max(0, P->n_cu / kQueues / 2);
  NW_TRY(fast2d_steps_queue_items(P->gf, d.order, d.absorbing, P->rows[0], P->tblock,
                                  d.n_shots, P->spw[0], split_tail, items, off, &per_cu));
  NW_HIP(hipDeviceSynchronize());
  if ((int64_t)items.size() > P->wq_cap) {
    if (P->wq_items) NW_HIP(hipFree(P->wq_items));
    P->wq_items = nullptr;
    P->wq_cap = 0;
    NW_HIP(hipMalloc((void**)&P->wq_items, items.size() * sizeof(int4)));
    P->wq_cap = (int64_t)items.size();
  }
  NW_HIP(hipMemcpy(P->wq_items, items.data(), items.size() * sizeof(int4),
                   hipMemcpyHostToDevice));
  if (!P->wq.ctr) {
    NW_HIP(hipMalloc((void**)&P->wq.ctr, (kQueues + 1) * sizeof(int)));
    NW_HIP(hipMemset(P->wq.ctr, 0, (kQueues + 1) * sizeof(int)));
    NW_HIP(hipDeviceSynchronize());
  }
  P->wq.items = P->wq_items;
  for (int j = 0; j <= kQueues; ++j) P->wq.off[j] = off[j];
  P->wq.grid = (int)std::min<int64_t>((int64_t)P->n_cu * per_cu, (int64_t)items.size());
  if (P->wq.grid < 1) return NW_OK;
  P->wq_key[0] = P->rows[0];
  P->wq_key[1] = P->tblock;
  P->wq_key[2] = P->spw[0];
  return NW_OK;
}

int32_t nw_plan_set_tiles(nw_plan* P, int32_t kind, int32_t rows, int32_t shots_per_group) {
  if (!P || kind < 0 || kind > 1) return fail(NW_ERR_ARG, "nw_plan_set_tiles: bad plan or kind");
  drop_graphs(P);
  if (rows == 0) {
    P->rows[kind] = 0;
    P->spw[kind] = 1;
    P->lds_tiles[kind] = 0;
    return kind == 0 ? build_queue(P) : NW_OK;
  }
  if (!fast2d_supported(P->gf, P->d.absorbing, P->d.dtype))
    return fail(NW_ERR_UNSUPPORTED, "nw_plan_set_tiles: the fast 2-D kernels do not cover this plan");
  if (rows != 4 && rows != 8 && rows != 16 && rows != 32 && rows != 64)
    return fail(NW_ERR_ARG, "nw_plan_set_tiles: rows must be 4, 8, 16, 32 or 64");
  if (shots_per_group < 1 || shots_per_group > P->d.n_shots)
    return fail(NW_ERR_ARG, "nw_plan_set_tiles: shots_per_group must be in [1, n_shots]");
  P->rows[kind] = rows;
  P->spw[kind] = shots_per_group;
  if (rows < 8 || rows > 32) P->lds_tiles[kind] = 0;
  return kind == 0 ? build_queue(P) : NW_OK;
}

int32_t nw_plan_get_tiles(const nw_plan* P, int32_t kind, int32_t* rows, int32_t* shots_per_group) {
  if (!P || kind < 0 || kind > 1 || !rows || !shots_per_group)
    return fail(NW_ERR_ARG, "nw_plan_get_tiles: bad argument");
  *rows = P->rows[kind];
  *shots_per_group = P->spw[kind];
  return NW_OK;
}

int32_t nw_plan_set_option(nw_plan* P, int32_t option, int32_t value) {
  if (!P) return fail(NW_ERR_ARG, "nw_plan_set_option: null plan");
  if (option == NW_OPT_GRAPH) {  // keeps the graphs captured so far
    P->graph_mode = value != 0;
    return NW_OK;
  }
  drop_graphs(P);  // captured graphs bake the kernels the other options select
  if (option == NW_OPT_XCD_MAP) {
    P->xcd_map = value != 0;
    return NW_OK;
  }
  if (option == NW_OPT_ROWS_PER_ITER) {
    if (value != 1 && value != 2) return fail(NW_ERR_ARG, "rows per iteration must be 1 or 2");
    P->rows_per_iter = value;
    return NW_OK;
  }
  if (option == NW_OPT_LDS_TILES_STEP || option == NW_OPT_LDS_TILES_ADJOINT) {
    const int k = option == NW_OPT_LDS_TILES_STEP ? 0 : 1;
    if (value && (P->rows[k] < 8 || P->rows[k] > 32))
      return fail(NW_ERR_ARG, "LDS tiles need rows of 8, 16 or 32");
    P->lds_tiles[k] = value != 0;
    return NW_OK;
  }
  if (option == NW_OPT_KEEP_LAP) {
    P->keep_lap_opt = value != 0;
    return NW_OK;
  }
  if (option == NW_OPT_FAST_3D) {
    if (value && !fast3d_supported(P->gf, P->d.absorbing, P->d.dtype))
      return fail(NW_ERR_UNSUPPORTED, "the fast 3-D kernels do not cover this plan");
    P->fast3d = value != 0;
    return NW_OK;
  }
  if (option == NW_OPT_FAST_F64) {
    if (value && !fast2d64_supported(P->gd, P->d.absorbing, P->d.dtype))
      return fail(NW_ERR_UNSUPPORTED, "the tiled float64 step does not cover this plan");
    P->fast64 = value != 0;
    return NW_OK;
  }
  if (option == NW_OPT_TIME_BLOCK || option == NW_OPT_TIME_BLOCK_ADJOINT) {
    if (value < 1 || value > kMaxDepth)
      return fail(NW_ERR_ARG, "time block depth must be 1.." + std::to_string(kMaxDepth));
    if (value > 1 && !P->tb_state[0]) {
      if (P->d.dtype != NW_F32 || P->d.absorbing == NW_PML ||
          !fast2d_supported(P->gf, P->d.absorbing, P->d.dtype))
        return fail(NW_ERR_UNSUPPORTED,
                    "time blocking needs the fast 2-D float32 kernels without a PML");
      const int64_t taps = (int64_t)(kMaxDepth - 1) * P->d.n_shots *
                           std::max(1, P->d.rec_max * P->d.rec_taps);
      const int64_t inj = (int64_t)kMaxDepth * P->d.n_shots * std::max(1, P->d.inj_max);
      for (int k = 0; k < 2; ++k)
        NW_HIP(hipMalloc((void**)&P->tb_state[k], (size_t)P->shot_bytes));
      NW_HIP(hipMalloc((void**)&P->tapbuf, (size_t)taps * sizeof(float)));
      NW_HIP(hipMalloc((void**)&P->injv, (size_t)inj * sizeof(float)));
      NW_TRY(build_tap_rows(P));
    }
    if (value > 1 && option == NW_OPT_TIME_BLOCK_ADJOINT && !P->tb_adj[0])
      for (int k = 0; k < 2; ++k)
        NW_HIP(hipMalloc((void**)&P->tb_adj[k], (size_t)P->shot_bytes));
    (option == NW_OPT_TIME_BLOCK ? P->tblock : P->tblock_adj) = value;
    return option == NW_OPT_TIME_BLOCK ? build_queue(P) : NW_OK;
  }
  if (option == NW_OPT_WORK_QUEUE) {
    if (value < 0 || value > 2) return fail(NW_ERR_ARG, "work queue mode must be 0, 1 or 2");
    P->work_queue = value;
    return build_queue(P);
  }
  return fail(NW_ERR_ARG, "nw_plan_set_option: unknown option");
}

int32_t nw_plan_get_option(const nw_plan* P, int32_t option, int32_t* value) {
  if (!P || !value) return fail(NW_ERR_ARG, "nw_plan_get_option: null argument");
  if (option == NW_OPT_GRAPH) {
    *value = P->graph_mode;
    return NW_OK;
  }
  if (option == NW_OPT_GRAPH_COUNT) {
    *value = (int32_t)P->graphs.size();
    return NW_OK;
  }
  if (option == NW_OPT_XCD_MAP) {
    *value = P->xcd_map;
    return NW_OK;
  }
  if (option == NW_OPT_ROWS_PER_ITER) {
    *value = P->rows_per_iter;
    return NW_OK;
  }
  if (option == NW_OPT_LDS_TILES_STEP || option == NW_OPT_LDS_TILES_ADJOINT) {
    *value = P->lds_tiles[option == NW_OPT_LDS_TILES_STEP ? 0 : 1];
    return NW_OK;
  }
  if (option == NW_OPT_TIME_BLOCK) {
    *value = P->tblock;
    return NW_OK;
  }
  if (option == NW_OPT_TIME_BLOCK_ADJOINT) {
    *value = P->tblock_adj;
    return NW_OK;
  }
  if (option == NW_OPT_KEEP_LAP) {
    *value = keep_lap(P) ? 1 : 0;
    return NW_OK;
  }
  if (option == NW_OPT_FAST_3D) {
    *value = P->fast3d;
    return NW_OK;
  }
  if (option == NW_OPT_FAST_F64) {
    *value = P->fast64;
    return NW_OK;
  }
  if (option == NW_OPT_WORK_QUEUE) {
    *value = P->work_queue;
    return NW_OK;
  }
  if (option == NW_OPT_WORK_QUEUE_ITEMS) {
    *value = P->wq_key[0] ? P->wq.off[kQueues] : 0;
    return NW_OK;
  }
  return fail(NW_ERR_ARG, "nw_plan_get_option: unknown option");
}

int64_t nw_ckpt_slot_bytes(const nw_plan* P) {
  return 2 * P->shot_bytes + P->face_bytes[0] + P->face_bytes[1] + P->face_bytes[2];
}

int32_t nw_face_layout(const nw_plan* P, int32_t axis, int64_t* elems, int64_t* row_pitch) {
  if (!P || axis < 0 || axis > 2 || !elems || !row_pitch)
    return fail(NW_ERR_ARG, "nw_face_layout: bad plan, axis or output pointer");
  const int64_t es = P->d.dtype == NW_F64 ? 8 : 4;
  *elems = P->face_bytes[axis] / es;
  *row_pitch = axis == 2 ? P->gf.FP2 : P->gf.D[2];
  return NW_OK;
}

int64_t nw_keep_slot_bytes(const nw_plan* P) {
  // a slot of kept Laplacians holds no face fields (the reverse step reads none)
  if (nw::keep_lap(P)) return P->shot_bytes;
  return P->shot_bytes + P->face_bytes[0] + P->face_bytes[1] + P->face_bytes[2];
}

int32_t nw_model(nw_plan* P, const void* c, void* m, void* ma, void* stream) {
  if (!P || !c || !m) return fail(NW_ERR_ARG, "nw_model: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t* D = P->d.dims;
  const dim3 blk = block_for(D);
  const dim3 grd = grid_for(D, 1, blk);
#define NW_MODEL(T, A) \
  hipLaunchKernelGGL((k_model<T, A>), grd, blk, 0, st, geo<T>(P), (const T*)c, (T*)m, (T*)ma)
  if (P->d.dtype == NW_F32) {
    switch (P->d.absorbing) {
      case NW_NONE: NW_MODEL(float, NW_NONE); break;
      case NW_SPONGE: NW_MODEL(float, NW_SPONGE); break;
      case NW_CERJAN: NW_MODEL(float, NW_CERJAN); break;
      default: NW_MODEL(float, NW_PML); break;
    }
  } else {
    switch (P->d.absorbing) {
      case NW_NONE: NW_MODEL(double, NW_NONE); break;
      case NW_SPONGE: NW_MODEL(double, NW_SPONGE); break;
      case NW_CERJAN: NW_MODEL(double, NW_CERJAN); break;
      default: NW_MODEL(double, NW_PML); break;
    }
  }
#undef NW_MODEL
  return check_launch("k_model");
}

int32_t nw_forward(nw_plan* P, const void* m, void** state, void** psi, int32_t n_begin,
                   int32_t n_end, void* traces, void* ckpt, int32_t ckpt_every, void* keep,
                   int32_t keep_begin, int32_t first_step, void* stream) {
  if (!P || !m || !state || !state[0] || !state[1]) return fail(NW_ERR_ARG, "nw_forward: null argument");
  if (n_begin < 0 || n_end < n_begin || n_end > P->d.n_steps)
    return fail(NW_ERR_ARG, "nw_forward: step range outside [0, n_steps]");
  if (P->d.absorbing == NW_PML && !psi) return fail(NW_ERR_ARG, "nw_forward: PML needs psi buffers");
  if (keep && keep_begin != n_begin) return fail(NW_ERR_ARG, "nw_forward: keep_begin must equal n_begin");
  if (ckpt && ckpt_every < 1) return fail(NW_ERR_ARG, "nw_forward: ckpt_every must be >= 1");
  const int npsi = P->d.absorbing == NW_PML ? 2 * P->d.ndim : 0;
  if (!aligned16(m) || !all_aligned16(state, 2) || (npsi && !all_aligned16(psi, npsi)) ||
      !aligned16(ckpt) || !aligned16(keep))
    return fail(NW_ERR_ARG, "nw_forward: device buffers must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  std::vector<uintptr_t> key = {0, (uintptr_t)m, (uintptr_t)n_begin, (uintptr_t)n_end,
                                (uintptr_t)traces, (uintptr_t)ckpt, (uintptr_t)ckpt_every,
                                (uintptr_t)keep, (uintptr_t)keep_begin, (uintptr_t)first_step};
  return run_captured(P, key, {state, psi}, {2, npsi}, st, [&](hipStream_t s2) {
    if (P->d.dtype == NW_F32)
      return forward_impl<float>(P, (const float*)m, state, psi, n_begin, n_end, (float*)traces,
                                 (float*)ckpt, ckpt_every, (float*)keep, keep_begin, first_step,
                                 s2);
    return forward_impl<double>(P, (const double*)m, state, psi, n_begin, n_end,
                                (double*)traces, (double*)ckpt, ckpt_every, (double*)keep,
                                keep_begin, first_step, s2);
  });
}

int32_t nw_adjoint(nw_plan* P, const void* m, const void* ma, void** lam, void** psi_bar,
                   int32_t n_begin, int32_t n_end, const void* keep, int32_t keep_begin,
                   const void* residual, int32_t inject_end, double* grad_m, int32_t first_step,
                   void* stream) {
  if (!P || !ma || !lam || !lam[0] || !lam[1]) return fail(NW_ERR_ARG, "nw_adjoint: null argument");
  if (n_begin < 0 || n_end < n_begin || n_end > P->d.n_steps)
    return fail(NW_ERR_ARG, "nw_adjoint: step range outside [0, n_steps]");
  if (P->d.absorbing == NW_PML && !psi_bar) return fail(NW_ERR_ARG, "nw_adjoint: PML needs psi_bar");
  if (keep && keep_begin > n_begin) return fail(NW_ERR_ARG, "nw_adjoint: keep must start <= n_begin");
  if (grad_m && !keep) return fail(NW_ERR_ARG, "nw_adjoint: grad_m needs the kept wavefields");
  const int npsi = P->d.absorbing == NW_PML ? 2 * P->d.ndim : 0;
  if (!aligned16(ma) || !all_aligned16(lam, 2) || (npsi && !all_aligned16(psi_bar, npsi)) ||
      !aligned16(keep) || !aligned16(grad_m))
    return fail(NW_ERR_ARG, "nw_adjoint: device buffers must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  std::vector<uintptr_t> key = {1, (uintptr_t)ma, (uintptr_t)n_begin, (uintptr_t)n_end,
                                (uintptr_t)keep, (uintptr_t)keep_begin, (uintptr_t)residual,
                                (uintptr_t)inject_end, (uintptr_t)grad_m, (uintptr_t)first_step};
  return run_captured(P, key, {lam, psi_bar}, {2, npsi}, st, [&](hipStream_t s2) {
    if (P->d.dtype == NW_F32)
      return adjoint_impl<float>(P, (const float*)m, (const float*)ma, lam, psi_bar, n_begin,
                                 n_end, (const float*)keep, keep_begin, (const float*)residual,
                                 inject_end, grad_m, first_step, s2);
    return adjoint_impl<double>(P, (const double*)m, (const double*)ma, lam, psi_bar, n_begin,
                                n_end, (const double*)keep, keep_begin, (const double*)residual,
                                inject_end, grad_m, first_step, s2);
  });
}

int32_t nw_model_grad(nw_plan* P, const void* c, const double* grad_m, void* grad_c, void* stream) {
  if (!P || !c || !grad_m || !grad_c) return fail(NW_ERR_ARG, "nw_model_grad: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nf = P->d.field_dims[0] * P->d.field_dims[1] * P->d.field_dims[2];
  const dim3 grd((unsigned)((nf + 255) / 256));
  const int64_t slices = nw_grad_slices(P);
  if (P->d.dtype == NW_F32)
    hipLaunchKernelGGL(k_model_grad<float>, grd, dim3(256), 0, st, P->gf, (const float*)c, grad_m,
                       (float*)grad_c, slices);
  else
    hipLaunchKernelGGL(k_model_grad<double>, grd, dim3(256), 0, st, P->gd, (const double*)c,
                       grad_m, (double*)grad_c, slices);
  return check_launch("k_model_grad");
}

int32_t nw_ic_state(nw_plan* P, const void* m, const void* p0, void* p_prev, void* p, void* stream) {
  if (!P || !m || !p0 || !p_prev || !p) return fail(NW_ERR_ARG, "nw_ic_state: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t* D = P->d.dims;
  const dim3 blk = block_for(D);
  const dim3 grd = grid_for(D, P->d.n_shots, blk);
  const bool o4 = P->d.order == 4;
  if (P->d.dtype == NW_F32) {
    if (o4) hipLaunchKernelGGL((k_ic_state<float, 2>), grd, blk, 0, st, P->gf, (const float*)m, (const float*)p0, (float*)p_prev, (float*)p);
    else hipLaunchKernelGGL((k_ic_state<float, 1>), grd, blk, 0, st, P->gf, (const float*)m, (const float*)p0, (float*)p_prev, (float*)p);
  } else {
    if (o4) hipLaunchKernelGGL((k_ic_state<double, 2>), grd, blk, 0, st, P->gd, (const double*)m, (const double*)p0, (double*)p_prev, (double*)p);
    else hipLaunchKernelGGL((k_ic_state<double, 1>), grd, blk, 0, st, P->gd, (const double*)m, (const double*)p0, (double*)p_prev, (double*)p);
  }
  return check_launch("k_ic_state");
}

int32_t nw_ic_grad(nw_plan* P, const void* m, const void* lam0, const void* pbar, void* grad_p0,
                   void* stream) {
  if (!P || !m || !lam0 || !pbar || !grad_p0) return fail(NW_ERR_ARG, "nw_ic_grad: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t* D = P->d.dims;
  const dim3 blk = block_for(D);
  const dim3 grd = grid_for(D, P->d.n_shots, blk);
  const bool o4 = P->d.order == 4;
  if (P->d.dtype == NW_F32) {
    if (o4) hipLaunchKernelGGL((k_ic_grad<float, 2>), grd, blk, 0, st, P->gf, (const float*)m, (const float*)lam0, (const float*)pbar, (float*)grad_p0);
    else hipLaunchKernelGGL((k_ic_grad<float, 1>), grd, blk, 0, st, P->gf, (const float*)m, (const float*)lam0, (const float*)pbar, (float*)grad_p0);
  } else {
    if (o4) hipLaunchKernelGGL((k_ic_grad<double, 2>), grd, blk, 0, st, P->gd, (const double*)m, (const double*)lam0, (const double*)pbar, (double*)grad_p0);
    else hipLaunchKernelGGL((k_ic_grad<double, 1>), grd, blk, 0, st, P->gd, (const double*)m, (const double*)lam0, (const double*)pbar, (double*)grad_p0);
  }
  return check_launch("k_ic_grad");
}

int32_t nw_prev_adjoint(nw_plan* P, const void* lam1, void* pbar, void* stream) {
  if (!P || !lam1 || !pbar) return fail(NW_ERR_ARG, "nw_prev_adjoint: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t* D = P->d.dims;
  const dim3 blk = block_for(D);
  const dim3 grd = grid_for(D, P->d.n_shots, blk);
#define NW_PA(T, A) \
  hipLaunchKernelGGL((k_prev_adjoint<T, A>), grd, blk, 0, st, geo<T>(P), (const T*)lam1, (T*)pbar)
  if (P->d.dtype == NW_F32) {
    switch (P->d.absorbing) {
      case NW_NONE: NW_PA(float, NW_NONE); break;
      case NW_SPONGE: NW_PA(float, NW_SPONGE); break;
      case NW_CERJAN: NW_PA(float, NW_CERJAN); break;
      default: NW_PA(float, NW_PML); break;
    }
  } else {
    switch (P->d.absorbing) {
      case NW_NONE: NW_PA(double, NW_NONE); break;
      case NW_SPONGE: NW_PA(double, NW_SPONGE); break;
      case NW_CERJAN: NW_PA(double, NW_CERJAN); break;
      default: NW_PA(double, NW_PML); break;
    }
  }
#undef NW_PA
  return check_launch("k_prev_adjoint");
}

int32_t nw_laplacian(nw_plan* P, const void* p, void* out, int64_t count, void* stream) {
  if (!P || !p || !out || count < 1) return fail(NW_ERR_ARG, "nw_laplacian: bad argument");
  if (P->d.dims[0] * count > 65535) return fail(NW_ERR_ARG, "nw_laplacian: too many grids");
  hipStream_t st = (hipStream_t)stream;
  const int64_t* D = P->d.dims;
  const dim3 blk = block_for(D);
  const dim3 grd = grid_for(D, count, blk);
  const bool o4 = P->d.order == 4;
  if (P->d.dtype == NW_F32) {
    if (o4) hipLaunchKernelGGL((k_laplacian<float, 2>), grd, blk, 0, st, P->gf, (const float*)p, (float*)out);
    else hipLaunchKernelGGL((k_laplacian<float, 1>), grd, blk, 0, st, P->gf, (const float*)p, (float*)out);
  } else {
    if (o4) hipLaunchKernelGGL((k_laplacian<double, 2>), grd, blk, 0, st, P->gd, (const double*)p, (double*)out);
    else hipLaunchKernelGGL((k_laplacian<double, 1>), grd, blk, 0, st, P->gd, (const double*)p, (double*)out);
  }
  return check_launch("k_laplacian");
}

int32_t nw_clamp(nw_plan* P, void* p, const int32_t* cells, const void* values, int32_t n_cells,
                 void* stream) {
  if (!P || !p || !cells || !values || n_cells < 0) return fail(NW_ERR_ARG, "nw_clamp: bad argument");
  if (n_cells == 0) return NW_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t cellsz = P->gf.cells;
  const dim3 grd((n_cells + 255) / 256, P->d.n_shots);
  if (P->d.dtype == NW_F32)
    hipLaunchKernelGGL(k_clamp<float>, grd, dim3(256), 0, st, cellsz, (float*)p, cells, (const float*)values, n_cells);
  else
    hipLaunchKernelGGL(k_clamp<double>, grd, dim3(256), 0, st, cellsz, (double*)p, cells, (const double*)values, n_cells);
  return check_launch("k_clamp");
}

}  // extern "C"
