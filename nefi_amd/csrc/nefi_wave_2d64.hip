// LDS-tiled 2-D and 3-D float64 forward steps of libnefi_wave (sponge, Cerjan and no absorbing
// layer; order 2/4): the step that float64 data generation runs, nefi's data generators calling
// the operator at twice the resolution in float64 (bench/base.py:79-90). The 2-D tile is
// described here; the 3-D one with its kernel below.
//
// Mapping: a workgroup of 4 wavefronts owns TY rows x 128 columns of one shot; lane l owns the
// two columns x0 + 2l, x0 + 2l + 1 (one 16-byte vector, a wavefront covers 1 KiB per row) and
// wavefront w computes rows w*TY/4 .. (w+1)*TY/4 - 1. The tile of p^n with R halo rows and 2
// margin columns on each side is loaded into LDS in one round trip; every stencil operand then
// comes from LDS, and p^{n-1}, m are read once per cell. Tiles are numbered column chunk, row
// strip, shot, with each XCD taking a contiguous run (so the strips above and below a tile,
// whose rows form its halo, share that XCD's L2).
//
// Arithmetic: the generic kernel's expressions (nefi_wave.hip, k_step), term for term and in
// the same order: centre term, then axis 1 and axis 2 contributions each summed before being
// added; damping from the row and column profile values; rhs = (2 p - b p_prev) + m lap.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nefi_wave.h"
#include "nefi_wave_internal.h"

namespace nw {

namespace {

constexpr int kXcd64 = 8;
constexpr int kLW64 = 128 + 4;  // doubles per LDS row: 2 margin columns on each side

__device__ __forceinline__ double2 ld2d(const double* p) {
  return *reinterpret_cast<const double2*>(p);
}
__device__ __forceinline__ void st2d(double* p, const double2& v) {
  *reinterpret_cast<double2*>(p) = v;
}

template <int TY, int R>
struct Shape64 {
  static constexpr int rows = TY + 2 * R;
  static constexpr int n2 = rows * (kLW64 / 2);  // double2 elements of one tile
  static constexpr int per_thread = (n2 + 255) / 256;
  static constexpr int my_rows = TY / 4;
};

// one cell's stencil from its LDS neighbours, in the generic kernel's order
template <int R>
__device__ __forceinline__ double lap64(const Geo<double>& g, double c, double ym1, double yp1,
                                        double ym2, double yp2, double xm1, double xp1,
                                        double xm2, double xp2) {
  double s = g.cc * c;
  double acc = g.c1[1] * (ym1 + yp1);
  if (R == 2) acc += g.c2[1] * (ym2 + yp2);
  s += acc;
  acc = g.c1[2] * (xm1 + xp1);
  if (R == 2) acc += g.c2[2] * (xm2 + xp2);
  s += acc;
  return s;
}

template <int ABS>
__device__ __forceinline__ void damp64(double hy, double hx, double& a, double& b, double& tap) {
  a = 1.0;
  b = 1.0;
  tap = 1.0;
  if (ABS == NW_SPONGE) {
    double half = 0.0;
    half += hy;
    half += hx;
    a = 1.0 / (1.0 + half);
    b = 1.0 - half;
  } else if (ABS == NW_CERJAN) {
    tap = hy < tap ? hy : tap;
    tap = hx < tap ? hx : tap;
    a = tap;
  }
}

template <int ABS>
__device__ __forceinline__ double cell64(double pc, double prev_raw, double m, double lap,
                                         double hy, double hx, int taper_prev, double src) {
  double a, b, tap;
  damp64<ABS>(hy, hx, a, b, tap);
  const double prev = (ABS == NW_CERJAN && taper_prev) ? tap * prev_raw : prev_raw;
  double rhs = (2.0 * pc - b * prev) + m * lap;
  rhs += src;
  return ABS == NW_NONE ? rhs : a * rhs;
}

template <int R, int ABS, int TY>
__global__ void __launch_bounds__(256) k2d64_step(Geo<double> g, StepIO<double> io, int shots,
                                                  int xcd_map) {
  using S = Shape64<TY, R>;
  __shared__ __align__(16) double lp[S::rows * kLW64];
  const int D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int nX = (D2 + 127) / 128, nY = (D1 + TY - 1) / TY;
  const int T = nX * nY * shots;
  const int per = (T + kXcd64 - 1) / kXcd64;
  const int bi = blockIdx.x;
  const int t = xcd_map ? (bi % kXcd64) * per + bi / kXcd64 : bi;
  if (t >= T) return;  // uniform over the workgroup
  const int x0 = (t % nX) * 128, y0 = ((t / nX) % nY) * TY, s = t / (nX * nY);
  const int64_t base = (int64_t)s * g.cells;
  const double* P = io.p + base;
  double2 v[S::per_thread];
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int row = i / (kLW64 / 2), c2 = i % (kLW64 / 2);
    const int gy = y0 - R + row, gx = x0 - 2 + 2 * c2;
    const bool ok = i < S::n2 && gy >= 0 && gy < D1 && gx >= 0 && gx < D2;
    v[k] = ok ? ld2d(P + (gy * P2 + gx)) : make_double2(0.0, 0.0);
  }
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    if (i < S::n2) st2d(lp + 2 * i, v[k]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = x0 + 2 * lane;
  if (x >= D2) return;  // no barrier follows
  const double hx0 = ABS != NW_NONE ? g.damp[2][x] : 0.0;
  const double hx1 = ABS != NW_NONE ? g.damp[2][x + 1] : 0.0;
  const double w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.0;
#pragma unroll
  for (int k = 0; k < S::my_rows; ++k) {
    const int y = y0 + wave * S::my_rows + k;
    if (y >= D1) break;
    const double* row = lp + (wave * S::my_rows + k + R) * kLW64 + 2 * lane;  // column x - 2
    const double2 l = ld2d(row), c = ld2d(row + 2), r = ld2d(row + 4);
    const double2 ym1 = ld2d(row - kLW64 + 2), yp1 = ld2d(row + kLW64 + 2);
    const double2 ym2 = R == 2 ? ld2d(row - 2 * kLW64 + 2) : make_double2(0.0, 0.0);
    const double2 yp2 = R == 2 ? ld2d(row + 2 * kLW64 + 2) : make_double2(0.0, 0.0);
    const double lap0 = lap64<R>(g, c.x, ym1.x, yp1.x, ym2.x, yp2.x, l.y, c.y, l.x, r.x);
    const double lap1 = lap64<R>(g, c.y, ym1.y, yp1.y, ym2.y, yp2.y, c.x, r.x, l.y, r.y);
    const int ro = y * P2 + x;
    const double2 pp = ld2d(io.p_prev + base + ro);
    const double2 mm = ld2d(io.m + ro);
    double src0 = 0.0, src1 = 0.0;
    for (int j = 0; j < io.src_taps; ++j) {
      const int d = io.src_idx[s * io.src_taps + j] - ro;
      if (d == 0) src0 += io.src_val[s * io.src_taps + j] * w_n;
      if (d == 1) src1 += io.src_val[s * io.src_taps + j] * w_n;
    }
    const double hy = ABS != NW_NONE ? g.damp[1][y] : 0.0;
    const int tpv = io.taper_prev;
    const double2 o = make_double2(cell64<ABS>(c.x, pp.x, mm.x, lap0, hy, hx0, tpv, src0),
                                   cell64<ABS>(c.y, pp.y, mm.y, lap1, hy, hx1, tpv, src1));
    st2d(io.p_next + base + ro, o);
    if (io.ck_prev) {
      st2d(io.ck_prev + base + ro, pp);
      st2d(io.ck_cur + base + ro, c);
    }
  }
}

// ------------------------------------------------------------------------------------------
// 3-D: 2.5-D tiles marching along the slowest axis
// ------------------------------------------------------------------------------------------
// A workgroup owns TY rows x 128 columns of one shot and a chunk of ZC planes of axis 0. The
// planes of p^n (R halo rows, 2 margin columns per side) sit in an LDS ring of 2R+1 planes;
// plane z is computed from the ring while plane z+R+1's loads are in flight in registers.
// Arithmetic: the generic kernel's expressions (centre term, then axes 0, 1, 2, each axis'
// contribution summed before being added; damping summed or minimised over the three
// profiles in axis order).
template <int TY, int R>
struct Ring64 {
  static constexpr int rows = TY + 2 * R;
  static constexpr int n2 = rows * (kLW64 / 2);
  static constexpr int per_thread = (n2 + 255) / 256;
  static constexpr int my_rows = TY / 4;
  static constexpr int slots = 2 * R + 1;
  static constexpr int plane = rows * kLW64;  // doubles of one LDS plane
};

template <int R>
__device__ __forceinline__ double lap3d64(const Geo<double>& g, double c, double zm1, double zp1,
                                          double zm2, double zp2, double ym1, double yp1,
                                          double ym2, double yp2, double xm1, double xp1,
                                          double xm2, double xp2) {
  double s = g.cc * c;
  double acc = g.c1[0] * (zm1 + zp1);
  if (R == 2) acc += g.c2[0] * (zm2 + zp2);
  s += acc;
  acc = g.c1[1] * (ym1 + yp1);
  if (R == 2) acc += g.c2[1] * (ym2 + yp2);
  s += acc;
  acc = g.c1[2] * (xm1 + xp1);
  if (R == 2) acc += g.c2[2] * (xm2 + xp2);
  s += acc;
  return s;
}

template <int ABS>
__device__ __forceinline__ double cell3d64(double pc, double prev_raw, double m, double lap,
                                           double hz, double hy, double hx, int taper_prev,
                                           double src) {
  double a = 1.0, b = 1.0, tap = 1.0;
  if (ABS == NW_SPONGE) {
    double half = 0.0;
    half += hz;
    half += hy;
    half += hx;
    a = 1.0 / (1.0 + half);
    b = 1.0 - half;
  } else if (ABS == NW_CERJAN) {
    tap = hz < tap ? hz : tap;
    tap = hy < tap ? hy : tap;
    tap = hx < tap ? hx : tap;
    a = tap;
  }
  const double prev = (ABS == NW_CERJAN && taper_prev) ? tap * prev_raw : prev_raw;
  double rhs = (2.0 * pc - b * prev) + m * lap;
  rhs += src;
  return ABS == NW_NONE ? rhs : a * rhs;
}

__device__ __forceinline__ int slot64(int z, int n) { return ((z % n) + n) % n; }

// loads of plane z of p^n's tile (zero outside the grid)
template <int TY, int R>
__device__ __forceinline__ void fetch64(const double* P, int z, int y0, int x0, int D0, int D1,
                                        int D2, int P2, int64_t plane, double2* v) {
  using S = Ring64<TY, R>;
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    const int row = i / (kLW64 / 2), c2 = i % (kLW64 / 2);
    const int gy = y0 - R + row, gx = x0 - 2 + 2 * c2;
    const bool ok = i < S::n2 && z >= 0 && z < D0 && gy >= 0 && gy < D1 && gx >= 0 && gx < D2;
    v[k] = ok ? ld2d(P + ((int64_t)z * plane + (gy * P2 + gx))) : make_double2(0.0, 0.0);
  }
}

template <int TY, int R>
__device__ __forceinline__ void put64(double* lds, const double2* v) {
  using S = Ring64<TY, R>;
#pragma unroll
  for (int k = 0; k < S::per_thread; ++k) {
    const int i = threadIdx.x + 256 * k;
    if (i < S::n2) st2d(lds + 2 * i, v[k]);
  }
}

template <int R, int ABS, int TY, int ZC>
__global__ void __launch_bounds__(256) k3d64_step(Geo<double> g, StepIO<double> io, int shots,
                                                  int xcd_map) {
  using S = Ring64<TY, R>;
  __shared__ __align__(16) double ring[S::slots * S::plane];
  const int D0 = (int)g.D[0], D1 = (int)g.D[1], D2 = (int)g.D[2], P2 = (int)g.P2;
  const int64_t plane = (int64_t)D1 * P2;
  const int nX = (D2 + 127) / 128, nY = (D1 + TY - 1) / TY, nZ = (D0 + ZC - 1) / ZC;
  const int T = nX * nY * nZ * shots;
  const int per = (T + kXcd64 - 1) / kXcd64;
  const int bi = blockIdx.x;
  const int t = xcd_map ? (bi % kXcd64) * per + bi / kXcd64 : bi;
  if (t >= T) return;  // uniform over the workgroup
  const int x0 = (t % nX) * 128, y0 = ((t / nX) % nY) * TY, z0 = ((t / (nX * nY)) % nZ) * ZC;
  const int s = t / (nX * nY * nZ);
  const int z1 = min(D0, z0 + ZC);
  const int64_t base = (int64_t)s * g.cells;
  const double* P = io.p + base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x = x0 + 2 * lane;
  const bool act = x < D2;
  const double hx0 = (ABS != NW_NONE && act) ? g.damp[2][x] : 0.0;
  const double hx1 = (ABS != NW_NONE && act) ? g.damp[2][x + 1] : 0.0;
  const double w_n = io.src_taps > 0 ? io.wavelet[io.n] : 0.0;
  double2 v[S::per_thread];
  for (int z = z0 - R; z < z0 + R; ++z) {
    fetch64<TY, R>(P, z, y0, x0, D0, D1, D2, P2, plane, v);
    put64<TY, R>(ring + slot64(z, S::slots) * S::plane, v);
  }
  fetch64<TY, R>(P, z0 + R, y0, x0, D0, D1, D2, P2, plane, v);
  for (int z = z0; z < z1; ++z) {
    __syncthreads();  // plane z-1 is computed: the slot of plane z-R-1 is free
    put64<TY, R>(ring + slot64(z + R, S::slots) * S::plane, v);
    __syncthreads();
    if (z + 1 < z1) fetch64<TY, R>(P, z + R + 1, y0, x0, D0, D1, D2, P2, plane, v);
    const double* pz = ring + slot64(z, S::slots) * S::plane;
    const double* pm1 = ring + slot64(z - 1, S::slots) * S::plane;
    const double* pp1 = ring + slot64(z + 1, S::slots) * S::plane;
    const double* pm2 = R == 2 ? ring + slot64(z - 2, S::slots) * S::plane : pz;
    const double* pp2 = R == 2 ? ring + slot64(z + 2, S::slots) * S::plane : pz;
    const double hz = ABS != NW_NONE ? g.damp[0][z] : 0.0;
#pragma unroll
    for (int k = 0; k < S::my_rows; ++k) {
      const int y = y0 + wave * S::my_rows + k;
      if (!act || y >= D1) continue;
      const int off = (wave * S::my_rows + k + R) * kLW64 + 2 * lane;  // column x - 2
      const double2 l = ld2d(pz + off), c = ld2d(pz + off + 2), r = ld2d(pz + off + 4);
      const double2 ym1 = ld2d(pz + off - kLW64 + 2), yp1 = ld2d(pz + off + kLW64 + 2);
      const double2 ym2 = R == 2 ? ld2d(pz + off - 2 * kLW64 + 2) : make_double2(0.0, 0.0);
      const double2 yp2 = R == 2 ? ld2d(pz + off + 2 * kLW64 + 2) : make_double2(0.0, 0.0);
      const double2 zm1 = ld2d(pm1 + off + 2), zp1 = ld2d(pp1 + off + 2);
      const double2 zm2 = R == 2 ? ld2d(pm2 + off + 2) : make_double2(0.0, 0.0);
      const double2 zp2 = R == 2 ? ld2d(pp2 + off + 2) : make_double2(0.0, 0.0);
      const double lap0 = lap3d64<R>(g, c.x, zm1.x, zp1.x, zm2.x, zp2.x, ym1.x, yp1.x, ym2.x,
                                     yp2.x, l.y, c.y, l.x, r.x);
      const double lap1 = lap3d64<R>(g, c.y, zm1.y, zp1.y, zm2.y, zp2.y, ym1.y, yp1.y, ym2.y,
                                     yp2.y, c.x, r.x, l.y, r.y);
      const int64_t ro = (int64_t)z * plane + (y * P2 + x);
      const double2 pp = ld2d(io.p_prev + base + ro);
      const double2 mm = ld2d(io.m + ro);
      double src0 = 0.0, src1 = 0.0;
      for (int j = 0; j < io.src_taps; ++j) {
        const int64_t d = io.src_idx[s * io.src_taps + j] - ro;
        if (d == 0) src0 += io.src_val[s * io.src_taps + j] * w_n;
        if (d == 1) src1 += io.src_val[s * io.src_taps + j] * w_n;
      }
      const double hy = ABS != NW_NONE ? g.damp[1][y] : 0.0;
      const int tpv = io.taper_prev;
      const double2 o =
          make_double2(cell3d64<ABS>(c.x, pp.x, mm.x, lap0, hz, hy, hx0, tpv, src0),
                       cell3d64<ABS>(c.y, pp.y, mm.y, lap1, hz, hy, hx1, tpv, src1));
      st2d(io.p_next + base + ro, o);
      if (io.ck_prev) {
        st2d(io.ck_prev + base + ro, pp);
        st2d(io.ck_cur + base + ro, c);
      }
    }
  }
}

template <int R, int ABS>
void launch3d64(const Geo<double>& g, const StepIO<double>& io, int shots, int xcd_map,
                hipStream_t st) {
  constexpr int TY = 4, ZC = 8;
  const int T = (int)((g.D[2] + 127) / 128) * (int)((g.D[1] + TY - 1) / TY) *
                (int)((g.D[0] + ZC - 1) / ZC) * shots;
  const unsigned blocks = (unsigned)(kXcd64 * ((T + kXcd64 - 1) / kXcd64));
  hipLaunchKernelGGL((k3d64_step<R, ABS, TY, ZC>), dim3(blocks), dim3(256), 0, st, g, io, shots,
                     xcd_map);
}

template <int R, int ABS>
void launch64(const Geo<double>& g, const StepIO<double>& io, int shots, int xcd_map,
              hipStream_t st) {
  if (g.nd == 3) {
    launch3d64<R, ABS>(g, io, shots, xcd_map, st);
    return;
  }
  constexpr int TY = 16;
  const int T = (int)((g.D[2] + 127) / 128) * (int)((g.D[1] + TY - 1) / TY) * shots;
  const unsigned blocks = (unsigned)(kXcd64 * ((T + kXcd64 - 1) / kXcd64));
  hipLaunchKernelGGL((k2d64_step<R, ABS, TY>), dim3(blocks), dim3(256), 0, st, g, io, shots,
                     xcd_map);
}

}  // namespace

bool fast2d64_supported(const Geo<double>& g, int absorbing, int dtype) {
  return (g.nd == 2 || g.nd == 3) && dtype == NW_F64 && absorbing != NW_PML &&
         g.D[2] % 2 == 0 && g.P2 % 2 == 0;
}

void fast2d64_step(const Geo<double>& g, int order, int absorbing, const StepIO<double>& io,
                   int shots, int xcd_map, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  switch (absorbing * 2 + (order == 4 ? 1 : 0)) {
    case NW_NONE * 2 + 0: launch64<1, NW_NONE>(g, io, shots, xcd_map, st); break;
    case NW_NONE * 2 + 1: launch64<2, NW_NONE>(g, io, shots, xcd_map, st); break;
    case NW_SPONGE * 2 + 0: launch64<1, NW_SPONGE>(g, io, shots, xcd_map, st); break;
    case NW_SPONGE * 2 + 1: launch64<2, NW_SPONGE>(g, io, shots, xcd_map, st); break;
    case NW_CERJAN * 2 + 0: launch64<1, NW_CERJAN>(g, io, shots, xcd_map, st); break;
    case NW_CERJAN * 2 + 1: launch64<2, NW_CERJAN>(g, io, shots, xcd_map, st); break;
    default: break;  // PML plans never select this kernel
  }
}

}  // namespace nw
