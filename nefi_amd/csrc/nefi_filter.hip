// Trace filters and the fused trace misfit: the C ABI of include/nefi_filter.h.
//
// The reference's trace filters (nefi/instances/wave_fwi/losses.py:29-50,
// nefi/instances/photoacoustic3d/sensor.py:24-38) multiply the zero-padded real FFT of each
// trace by a real, even response and crop. For outputs t < n that is the linear convolution of
// the trace with the symmetric kernel g[k] = h[|k|], |k| < n, h = irfft(H, N)[:n]; on any
// circular length L >= 2n - 1 the circular convolution with g (wrapped) gives the same first n
// samples. A plan stores the spectrum G of that wrapped kernel for a power of two L, real
// because g is even, in bit-reversed order and divided by L.
//
// One workgroup filters two traces at a time as the real and imaginary parts of one complex
// sequence: an in-place decimation-in-frequency FFT in LDS leaves the spectrum in bit-reversed
// order, the product with G (real) keeps the two traces apart, and an in-place decimation-in-time
// inverse FFT takes the bit-reversed spectrum back to natural order, so no permutation pass is
// needed. Radix-2 stages run in groups in registers, up to five per LDS round trip; the last
// forward stages, the product with G and the first inverse stages share one. The residual
// (x - x_sub) is formed in the first forward pass, which reads the traces from global memory; the
// mask, the output scale and the float64 sums of the masked squares are fused into the last
// inverse pass, which writes the outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "nefi_filter.h"
#include "nefi_wave.h"
#include "nefi_wave_internal.h"

struct nw_filter {
  int dtype;
  int64_t n;
  int64_t L;
  int log2L;
  void* spec;        // L reals: G in bit-reversed order, / L
  void* tw;          // L/2 complex twiddles exp(-2 pi i k / L)
  double* partial;   // two float64 sums per workgroup of the last apply
};

namespace nw {
namespace {

constexpr int kFilterThreads = 256;
constexpr int kFilterBlocks = 2048;  // workgroups of one apply (a loop over row pairs)
static_assert(2 * kFilterBlocks <= NW_FILTER_NOALLOC_PARTIALS,
              "the allocation-free apply's array must hold every workgroup's partial sums");
constexpr int kMinLog2 = 6;
// FFT lengths: 16384 float2 (+1/32 padding: 132 KiB of LDS) or 8192 double2 (+1/16: 136 KiB)
constexpr double kPi = 3.141592653589793238462643383279502884;

template <typename T>
struct Cplx;
template <>
struct Cplx<float> {
  using type = float2;
};
template <>
struct Cplx<double> {
  using type = double2;
};

template <typename T>
struct FilterIO {
  const T* x;
  const T* xs;    // subtracted input (or null)
  const T* mask;  // weights of x's shape (or null)
  T* y;           // output (or null)
  const T* spec;
  const typename Cplx<T>::type* tw;
  int64_t rows;
  int64_t n;
  T scale;
  double* partial;
};

template <typename V>
__device__ __forceinline__ V cmul(const V& a, const V& b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

template <typename V>
__device__ __forceinline__ V cmul_conj(const V& a, const V& b) {  // a * conj(b)
  V r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.y * b.x - a.x * b.y;
  return r;
}

template <typename V>
__device__ __forceinline__ V mul_neg_i(const V& a) {  // a * (-i)
  V r;
  r.x = a.y;
  r.y = -a.x;
  return r;
}

template <typename V>
__device__ __forceinline__ V mul_i(const V& a) {  // a * i
  V r;
  r.x = -a.y;
  r.y = a.x;
  return r;
}

// the two float64 sums of a workgroup (fixed order: lanes by shuffle tree, then wavefronts)
__device__ __forceinline__ void block_sums(double v, double m, double* out) {
  __shared__ double red[2][kFilterThreads / 64];
  for (int off = 32; off >= 1; off >>= 1) {
    v += __shfl_xor(v, off, 64);
    m += __shfl_xor(m, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = v;
    red[1][wave] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kFilterThreads / 64; ++w) {
      a += red[0][w];
      b += red[1][w];
    }
    out[0] = a;
    out[1] = b;
  }
}

// Stages of the forward transform, in passes: the top pass takes R_TOP stages, each middle pass
// R_MAX, and the merged pass (the last forward stages, the product with the spectrum and the
// first inverse stages, all in registers) R_LAST. Float32: R_LAST 5, the top pass up to 5 (half
// of its group is zero padding) and middle passes of 4; float64 (twice the registers per
// element): 4, 4 and 3. Every pass is one LDS round trip: 3 for L = 8192 in float32
// (4 + 4 + 5 stages) where one pass per stage would take 13.
template <typename T, int LOG2L>
struct FftPlan {
  static constexpr int R_LAST = sizeof(T) == 4 ? 5 : 4;
  static constexpr int R_TOP_MAX = R_LAST;
  static constexpr int R_MAX = sizeof(T) == 4 ? 4 : 3;
  static constexpr int REST = LOG2L - R_LAST;  // >= 1 (LOG2L >= 6)
  static constexpr int N_MID = REST <= R_TOP_MAX ? 0 : (REST - R_TOP_MAX + R_MAX - 1) / R_MAX;
  static constexpr int R_TOP = REST - R_MAX * N_MID;
  static_assert(REST >= 1 && R_TOP >= 1 && R_TOP <= R_TOP_MAX, "pass plan");
  // workgroups of 4 wavefronts the LDS admits per CU (160 KiB), i.e. wavefronts per SIMD: the
  // register cap asks for no more than 2 (256 VGPRs)
  static constexpr int LDS_BYTES = ((1 << LOG2L) + ((1 << LOG2L) >> R_LAST)) * 2 * (int)sizeof(T);
  static constexpr int WAVES = (160 * 1024 / LDS_BYTES) >= 2 ? 2 : 1;
  // the top pass's weights held from the start of a pair (32 VGPRs at most)
  static constexpr int TOP_KG = (((1 << LOG2L) >> R_TOP) + kFilterThreads - 1) / kFilterThreads;
  static constexpr int MASK_REGS = 2 * TOP_KG * (1 << (R_TOP - 1)) * (int)sizeof(T) / 4;
  static constexpr bool MASK_AHEAD = MASK_REGS <= 32;
};

// LDS slot of sequence element e: one pad slot after every 2^R_LAST elements, so the merged
// pass's lanes (each on 2^R_LAST consecutive elements) start on different banks; the other
// passes' lanes walk consecutive elements.
template <typename T, int LOG2L>
__device__ __forceinline__ int slot(int e) {
  return e + (e >> FftPlan<T, LOG2L>::R_LAST);
}

// Slot distance of elements d apart when d is a multiple of the pad interval 2^R_LAST: then
// slot(e + d) = slot(e) + d + d / 2^R_LAST for every e, so a group's slots are its first slot
// plus compile-time offsets (immediates of the LDS instructions, no per-element arithmetic).
template <typename T, int LOG2L>
__device__ __forceinline__ constexpr int slot_step(int d) {
  return d + (d >> FftPlan<T, LOG2L>::R_LAST);
}

// exp(-2 pi i j / 32) for j < 16 (cos, -sin): the constant factors of a group's twiddles
constexpr double kRoot32[16][2] = {
    {1.0, 0.0},
    {0.98078528040323044913, -0.19509032201612826785},
    {0.92387953251128675613, -0.38268343236508977173},
    {0.83146961230254523708, -0.55557023301960222474},
    {0.70710678118654752440, -0.70710678118654752440},
    {0.55557023301960222474, -0.83146961230254523708},
    {0.38268343236508977173, -0.92387953251128675613},
    {0.19509032201612826785, -0.98078528040323044913},
    {0.0, -1.0},
    {-0.19509032201612826785, -0.98078528040323044913},
    {-0.38268343236508977173, -0.92387953251128675613},
    {-0.55557023301960222474, -0.83146961230254523708},
    {-0.70710678118654752440, -0.70710678118654752440},
    {-0.83146961230254523708, -0.55557023301960222474},
    {-0.92387953251128675613, -0.38268343236508977173},
    {-0.98078528040323044913, -0.19509032201612826785}};

// exp(-pi i m / 2^k) for the offsets m < 2^k <= 16 of a group stage (k <= 4)
template <typename V>
__device__ __forceinline__ V unit_root(int k, int m) {
  V r;
  const int j = m << (4 - k);
  r.x = kRoot32[j][0];
  r.y = kRoot32[j][1];
  return r;
}

// Raw buffer loads over one trace row: a resource of `bytes` bytes (stride 0) returns zero for a
// load past its end, which is the transform's zero padding, and a resource of 0 bytes stands for
// the missing second row of an odd row count. Word 3 selects 32-bit data (a zero data format
// would make every load read zero).
constexpr int kRsrcWord3 = 0x00020000;

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const T* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, bytes, kRsrcWord3);
}

template <typename T>
__device__ __forceinline__ T row_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Twiddle loads from the table through a buffer resource (SGPRs): a load is the lane's byte
// offset and nothing else, no 64-bit address arithmetic per twiddle.
template <typename T>
__device__ __forceinline__ typename Cplx<T>::type tw_load(__amdgpu_buffer_rsrc_t r, int voff) {
  using V = typename Cplx<T>::type;
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
  else
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
}

template <typename T, int LOG2L>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tw_rsrc(const typename Cplx<T>::type* tw) {
  return row_rsrc(tw, (int)sizeof(typename Cplx<T>::type) << (LOG2L - 1));  // L/2 entries
}

// The R consecutive radix-2 stages of one group, in registers. A group is the 2^R elements
// base + q s (q < 2^R) of a block of 2^(ls + R) elements, s = 2^ls: the stages of spans
// s 2^(R-1) down to s (forward) or s up to s 2^(R-1) (inverse) combine only them. Forward
// stages are decimation in frequency ((a + b, (a - b) w)), inverse stages decimation in time
// ((a + b conj(w), a - b conj(w))), with w = exp(-2 pi i j / (2h)) for the element's offset j in
// its half-block of span h. In a group j = off + m s (m < h / s), off = base mod s, so w is the
// table's exp(-2 pi i off / (2h)) times the constant exp(-pi i m s / h): one twiddle load per
// stage and group, none when the group starts its block (OFF false: off = 0).
template <typename T, int LOG2L, int R, bool INV, bool OFF>
__device__ __forceinline__ void group_stages(typename Cplx<T>::type* e, __amdgpu_buffer_rsrc_t twr,
                                             int ls, int off) {
  using V = typename Cplx<T>::type;
  constexpr int Q = 1 << R;
  constexpr int VS = sizeof(T) == 4 ? 3 : 4;  // log2 of a table entry's bytes
  static_assert(R <= 5, "the constant roots cover spans up to 16 in a group");
  V tb[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (OFF) {
      tb[k] = tw_load<T>(twr, off << (LOG2L - ls - k - 1 + VS));
    } else {
      tb[k].x = T(1);
      tb[k].y = T(0);
    }
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int k = INV ? t : R - 1 - t;  // the stage's span is s 2^k
    const int d = 1 << k;               // element distance in q
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (q & d) continue;
      const int m = q & (d - 1);
      // the constant factor exp(-pi i m / 2^k) is 1 for m = 0 and -i for m = 2^(k-1): those
      // products are exact swaps and sign flips (bit-identical to the general product)
      const bool unit = m == 0;
      const bool neg_i = m != 0 && (m << (4 - k)) == 8;
      const V a = e[q];
      V b = e[q + d];
      V w;
      if (OFF) {
        w = unit ? tb[k] : neg_i ? mul_neg_i(tb[k]) : cmul(tb[k], unit_root<V>(k, m));
      } else if (!unit && !neg_i) {
        w = unit_root<V>(k, m);
      }
      const bool trivial = !OFF && (unit || neg_i);
      if (INV) b = !trivial ? cmul_conj(b, w) : neg_i ? mul_i(b) : b;
      V sum, dif;
      sum.x = a.x + b.x;
      sum.y = a.y + b.y;
      dif.x = a.x - b.x;
      dif.y = a.y - b.y;
      e[q] = sum;
      e[q + d] = INV ? dif : !trivial ? cmul(dif, w) : neg_i ? mul_neg_i(dif) : dif;
    }
  }
}

// threadIdx.x as a value the compiler cannot see through: the group and twiddle index
// arithmetic of a pass is then computed where it is used, not hoisted out of the loop over trace
// pairs (where its many long-lived values would not fit in registers)
__device__ __forceinline__ int thread_id() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// One pass of R stages whose top span is 2^lh_top, LDS to LDS, CH groups in flight per thread
template <typename T, int LOG2L, int R, bool INV>
__device__ __forceinline__ void pass_lds(typename Cplx<T>::type* z,
                                         const typename Cplx<T>::type* tw, int lh_top) {
  using V = typename Cplx<T>::type;
  constexpr int L = 1 << LOG2L, Q = 1 << R, G = L / Q;
  constexpr int KG = (G + kFilterThreads - 1) / kFilterThreads;  // groups per thread
  constexpr int CH0 = sizeof(T) == 4 ? (R >= 4 ? 2 : 4) : (R >= 3 ? 1 : 2);
  constexpr int CH = KG < CH0 ? KG : CH0;  // groups loaded before the first store
  const int ls = lh_top - (R - 1);  // >= R_LAST: the passes above the merged one
  const int s = 1 << ls;
  const auto twr = tw_rsrc<T, LOG2L>(tw);
#pragma unroll
  for (int c0 = 0; c0 < KG; c0 += CH) {
    V e[CH][Q];
    int base[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int g = thread_id() + kFilterThreads * (c0 + c);
      base[c] = -1;
      if (G % kFilterThreads == 0 || g < G) {
        base[c] = ((g >> ls) << (ls + R)) + (g & (s - 1));
        const int zb = slot<T, LOG2L>(base[c]);
#pragma unroll
        for (int q = 0; q < Q; ++q) e[c][q] = z[zb + q * slot_step<T, LOG2L>(s)];
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      if (base[c] < 0) continue;
      group_stages<T, LOG2L, R, INV, true>(e[c], twr, ls, base[c] & (s - 1));
      const int zb = slot<T, LOG2L>(base[c]);
#pragma unroll
      for (int q = 0; q < Q; ++q) z[zb + q * slot_step<T, LOG2L>(s)] = e[c][q];
    }
  }
}

// The top passes' groups: group g < G = 2^ls holds t = g + q 2^ls (q < 2^R), so the lanes of a
// wavefront touch consecutive samples. Traces are at most L/2 long (L >= 2n - 1), so only the
// lower half q < 2^(R-1) of a group meets the traces: the upper half is zero on the way in and
// discarded on the way out.
template <typename T, int LOG2L, int R>
struct TopPass {
  static constexpr int Q = 1 << R, QH = Q / 2, G = (1 << LOG2L) / Q, ls = LOG2L - R;
  static constexpr int KG = (G + kFilterThreads - 1) / kFilterThreads;  // groups per thread
  __device__ static bool valid(int c) {
    return G % kFilterThreads == 0 || (int)thread_id() + kFilterThreads * c < G;
  }
};

// The first forward pass: its groups' elements come straight from the traces (the residual
// x - x_sub; rows r0, r0 + 1 as the real and imaginary parts, zero past the trace length). Lane
// offsets in VGPRs, q offsets in SGPRs, no branches: a group's loads are in flight together.
template <typename T, int LOG2L, int R, bool SUB>
__device__ __forceinline__ void pass_top_load_t(typename Cplx<T>::type* z,
                                                const typename Cplx<T>::type* tw,
                                                const FilterIO<T>& io, int64_t o0, int64_t o1,
                                                bool two) {
  using V = typename Cplx<T>::type;
  using P = TopPass<T, LOG2L, R>;
  const int bytes = (int)io.n * (int)sizeof(T);
  const auto x0 = row_rsrc(io.x + o0, bytes);
  const auto x1 = row_rsrc(io.x + o1, two ? bytes : 0);
  const auto s0 = row_rsrc(SUB ? io.xs + o0 : io.x, SUB ? bytes : 0);
  const auto s1 = row_rsrc(SUB ? io.xs + o1 : io.x, SUB && two ? bytes : 0);
#pragma unroll
  for (int c = 0; c < P::KG; ++c) {
    if (!P::valid(c)) continue;
    const int g = thread_id() + kFilterThreads * c;
    const int voff = g * (int)sizeof(T);
    T a[P::QH], b[P::QH], sa[P::QH], sb[P::QH];
#pragma unroll
    for (int q = 0; q < P::QH; ++q) {
      const int soff = (q << P::ls) * (int)sizeof(T);
      a[q] = row_load<T>(x0, voff, soff);
      b[q] = row_load<T>(x1, voff, soff);
      if (SUB) {
        sa[q] = row_load<T>(s0, voff, soff);
        sb[q] = row_load<T>(s1, voff, soff);
      }
    }
    // every load of the group issued before the first use (the scheduler would otherwise
    // interleave them with the arithmetic, one memory latency per sample)
    __builtin_amdgcn_sched_barrier(0);
    V e[P::Q];
#pragma unroll
    for (int q = 0; q < P::QH; ++q) {
      e[q].x = SUB ? a[q] - sa[q] : a[q];
      e[q].y = SUB ? b[q] - sb[q] : b[q];
      e[q + P::QH].x = T(0);
      e[q + P::QH].y = T(0);
    }
    group_stages<T, LOG2L, R, false, true>(e, tw_rsrc<T, LOG2L>(tw), P::ls, g);
    const int zb = slot<T, LOG2L>(g);
#pragma unroll
    for (int q = 0; q < P::Q; ++q) z[zb + q * slot_step<T, LOG2L>(1 << P::ls)] = e[q];
  }
}

template <typename T, int LOG2L, int R>
__device__ __forceinline__ void pass_top_load(typename Cplx<T>::type* z,
                                              const typename Cplx<T>::type* tw,
                                              const FilterIO<T>& io, int64_t o0, int64_t o1,
                                              bool two) {
  if (io.xs)
    pass_top_load_t<T, LOG2L, R, true>(z, tw, io, o0, o1, two);
  else
    pass_top_load_t<T, LOG2L, R, false>(z, tw, io, o0, o1, two);
}

// The weights of the samples the last inverse pass writes, loaded ahead (at the start of the
// pair) so their latency hides behind the LDS passes: the mask, 1 without one, and 0 past the
// trace length and for the missing second row, so the sums need no tests.
template <typename T, int LOG2L, int R>
struct TopMask {
  using P = TopPass<T, LOG2L, R>;
  T m[2][P::KG][P::QH];
  __device__ __forceinline__ void load(const FilterIO<T>& io, int64_t o0, int64_t o1, bool two) {
    const int n = (int)io.n, bytes = n * (int)sizeof(T);
    const auto m0 = row_rsrc(io.mask ? io.mask + o0 : io.x, io.mask ? bytes : 0);
    const auto m1 = row_rsrc(io.mask ? io.mask + o1 : io.x, io.mask && two ? bytes : 0);
#pragma unroll
    for (int c = 0; c < P::KG; ++c) {
      const int g = thread_id() + kFilterThreads * c;
#pragma unroll
      for (int q = 0; q < P::QH; ++q) {
        const int t = g + (q << P::ls);
        if (io.mask) {
          const int voff = g * (int)sizeof(T), soff = (q << P::ls) * (int)sizeof(T);
          m[0][c][q] = row_load<T>(m0, voff, soff);
          m[1][c][q] = row_load<T>(m1, voff, soff);
        } else {
          m[0][c][q] = t < n ? T(1) : T(0);
          m[1][c][q] = t < n && two ? T(1) : T(0);
        }
      }
    }
  }
};

// The last inverse pass: its groups' elements go straight to the outputs (mask, scale) and into
// the float64 sums of the masked squares.
template <typename T, int LOG2L, int R>
__device__ __forceinline__ void pass_top_store(const typename Cplx<T>::type* z,
                                               const typename Cplx<T>::type* tw,
                                               const FilterIO<T>& io, int64_t o0, int64_t o1,
                                               bool two, const TopMask<T, LOG2L, R>& mk,
                                               double& acc_v, double& acc_m) {
  using V = typename Cplx<T>::type;
  using P = TopPass<T, LOG2L, R>;
  const int n = (int)io.n;
#pragma unroll
  for (int c = 0; c < P::KG; ++c) {
    if (!P::valid(c)) continue;
    const int g = thread_id() + kFilterThreads * c;
    V e[P::Q];
    const int zb = slot<T, LOG2L>(g);
#pragma unroll
    for (int q = 0; q < P::Q; ++q) e[q] = z[zb + q * slot_step<T, LOG2L>(1 << P::ls)];
    group_stages<T, LOG2L, R, true, true>(e, tw_rsrc<T, LOG2L>(tw), P::ls, g);
#pragma unroll
    for (int q = 0; q < P::QH; ++q) {
      const int t = g + (q << P::ls);
      const V v = e[q];
      const T m0 = mk.m[0][c][q], m1 = mk.m[1][c][q];
      acc_v += (double)m0 * ((double)v.x * (double)v.x) + (double)m1 * ((double)v.y * (double)v.y);
      acc_m += (double)m0 + (double)m1;
      if (io.y && t < n) {
        io.y[o0 + t] = io.scale * (m0 * v.x);
        if (two) io.y[o1 + t] = io.scale * (m1 * v.y);
      }
    }
  }
}

// The last forward stages (spans 2^(R-1) .. 1: groups of 2^R consecutive elements), the product
// with the spectrum (real, at the bit-reversed positions the forward stages leave), and the
// first inverse stages on the same groups, in registers.
template <typename T, int LOG2L, int R>
__device__ __forceinline__ void pass_merged(typename Cplx<T>::type* z,
                                            const typename Cplx<T>::type* tw,
                                            const T* spec) {
  using V = typename Cplx<T>::type;
  constexpr int L = 1 << LOG2L, Q = 1 << R, G = L / Q;
  constexpr int KG = (G + kFilterThreads - 1) / kFilterThreads;
#pragma unroll 1
  for (int c = 0; c < KG; ++c) {
    const int g = thread_id() + kFilterThreads * c;
    if (G % kFilterThreads != 0 && g >= G) break;
    const int base = g << R;
    V e[Q];
    const int zb = slot<T, LOG2L>(base);  // base is a multiple of 2^R: its slots are zb + q
#pragma unroll
    for (int q = 0; q < Q; ++q) e[q] = z[zb + q];
    group_stages<T, LOG2L, R, false, false>(e, tw_rsrc<T, LOG2L>(tw), 0, 0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const T h = spec[base + q];
      e[q].x = e[q].x * h;
      e[q].y = e[q].y * h;
    }
    group_stages<T, LOG2L, R, true, false>(e, tw_rsrc<T, LOG2L>(tw), 0, 0);
#pragma unroll
    for (int q = 0; q < Q; ++q) z[zb + q] = e[q];
  }
}

template <typename T, int LOG2L>
__global__ void __launch_bounds__(kFilterThreads)
    __attribute__((amdgpu_waves_per_eu(FftPlan<T, LOG2L>::WAVES))) k_filter_fft(FilterIO<T> io) {
  using V = typename Cplx<T>::type;
  using F = FftPlan<T, LOG2L>;
  constexpr int L = 1 << LOG2L;
  __shared__ V z[L + (L >> F::R_LAST)];  // padded (slot)
  double acc_v = 0.0, acc_m = 0.0;
  const int64_t pairs = (io.rows + 1) / 2;
  for (int64_t p = blockIdx.x; p < pairs; p += gridDim.x) {
    const int64_t r0 = 2 * p;
    const bool two = r0 + 1 < io.rows;
    const int64_t o0 = r0 * io.n, o1 = two ? o0 + io.n : o0;
    __syncthreads();  // the previous pair's reads of z are done
    pass_top_load<T, LOG2L, F::R_TOP>(z, io.tw, io, o0, o1, two);
    TopMask<T, LOG2L, F::R_TOP> mk;
    if (F::MASK_AHEAD) mk.load(io, o0, o1, two);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < F::N_MID; ++q) {
      pass_lds<T, LOG2L, F::R_MAX, false>(z, io.tw, LOG2L - 1 - F::R_TOP - F::R_MAX * q);
      __syncthreads();
    }
    pass_merged<T, LOG2L, F::R_LAST>(z, io.tw, io.spec);
    __syncthreads();
#pragma unroll
    for (int q = F::N_MID - 1; q >= 0; --q) {
      pass_lds<T, LOG2L, F::R_MAX, true>(z, io.tw, LOG2L - 1 - F::R_TOP - F::R_MAX * q);
      __syncthreads();
    }
    if (!F::MASK_AHEAD) mk.load(io, o0, o1, two);
    pass_top_store<T, LOG2L, F::R_TOP>(z, io.tw, io, o0, o1, two, mk, acc_v, acc_m);
  }
  if (io.partial) block_sums(acc_v, acc_m, io.partial + 2 * blockIdx.x);
}

// the same passes without the transform (filter = 0)
template <typename T>
__global__ void __launch_bounds__(kFilterThreads) k_filter_ident(FilterIO<T> io) {
  const int64_t total = io.rows * io.n;
  double acc_v = 0.0, acc_m = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kFilterThreads + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * kFilterThreads) {
    T v = io.x[e];
    if (io.xs) v = v - io.xs[e];
    const T m = io.mask ? io.mask[e] : T(1);
    if (io.y) io.y[e] = io.scale * (m * v);
    acc_v += (double)m * ((double)v * (double)v);
    acc_m += (double)m;
  }
  if (io.partial) block_sums(acc_v, acc_m, io.partial + 2 * blockIdx.x);
}

// sums[k] = sum over workgroups of partial[2 w + k], in a fixed tree order
__global__ void __launch_bounds__(kFilterThreads) k_filter_sums(const double* partial, int blocks,
                                                               double* sums) {
  double v = 0.0, m = 0.0;
  for (int w = threadIdx.x; w < blocks; w += kFilterThreads) {
    v += partial[2 * w];
    m += partial[2 * w + 1];
  }
  block_sums(v, m, sums);
}

// host: in-place radix-2 complex FFT (forward, e^{-i}), double
void host_fft(std::vector<std::complex<double>>& a) {
  const size_t L = a.size();
  for (size_t i = 1, j = 0; i < L; ++i) {
    size_t bit = L >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= L; len <<= 1) {
    for (size_t k = 0; k < len / 2; ++k) {
      const double ang = -2.0 * kPi * (double)k / (double)len;
      const std::complex<double> w(std::cos(ang), std::sin(ang));
      for (size_t i = 0; i < L; i += len) {
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

size_t bitrev(size_t i, int bits) {
  size_t r = 0;
  for (int b = 0; b < bits; ++b) r |= ((i >> b) & 1) << (bits - 1 - b);
  return r;
}

template <typename T>
int upload(nw_filter* f, const std::vector<double>& spec, const std::vector<double>& tw) {
  std::vector<T> s(spec.begin(), spec.end()), t(tw.begin(), tw.end());
  if (hipMalloc(&f->spec, s.size() * sizeof(T)) != hipSuccess ||
      hipMalloc(&f->tw, t.size() * sizeof(T)) != hipSuccess ||
      hipMalloc((void**)&f->partial, 2 * kFilterBlocks * sizeof(double)) != hipSuccess)
    return fail(NW_ERR_HIP, "nw_filter_create: device allocation failed");
  if (hipMemcpy(f->spec, s.data(), s.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(f->tw, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return fail(NW_ERR_HIP, "nw_filter_create: copy to device failed");
  return NW_OK;
}

template <typename T, int LOG2L>
void launch_fft(const FilterIO<T>& io, int blocks, hipStream_t st) {
  hipLaunchKernelGGL((k_filter_fft<T, LOG2L>), dim3(blocks), dim3(kFilterThreads), 0, st, io);
}

template <typename T>
void dispatch_fft(int log2L, const FilterIO<T>& io, int blocks, hipStream_t st) {
  switch (log2L) {
    case 6: launch_fft<T, 6>(io, blocks, st); break;
    case 7: launch_fft<T, 7>(io, blocks, st); break;
    case 8: launch_fft<T, 8>(io, blocks, st); break;
    case 9: launch_fft<T, 9>(io, blocks, st); break;
    case 10: launch_fft<T, 10>(io, blocks, st); break;
    case 11: launch_fft<T, 11>(io, blocks, st); break;
    case 12: launch_fft<T, 12>(io, blocks, st); break;
    case 13: launch_fft<T, 13>(io, blocks, st); break;
    default:
      if constexpr (sizeof(T) == 4) launch_fft<T, 14>(io, blocks, st);
      break;
  }
}

// Partial sums of the allocation-free apply: one array per GPU, part of the code object, held
// until the process ends. Host code reaches it only through the per-device address that
// hipGetSymbolAddress returns (scratch_on_current_device), never through its name.
static __device__ __align__(256) double scratch[NW_FILTER_NOALLOC_PARTIALS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(NW_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(NW_ERR_UNSUPPORTED, "nw_filter_apply_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(NW_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return NW_OK;
}

// Workgroups one apply launches (each writes two partial sums).
int apply_blocks(const nw_filter* f, int64_t rows, int filter) {
  if (rows <= 0) return 0;
  if (filter) return (int)std::min<int64_t>((rows + 1) / 2, kFilterBlocks);
  const int64_t chunks = (rows * f->n + 4 * kFilterThreads - 1) / (4 * kFilterThreads);
  return (int)std::min<int64_t>(chunks, kFilterBlocks);
}

// `partial`: 2 doubles per workgroup for the sums (the plan's buffer or the per-GPU array)
template <typename T>
int apply_t(nw_filter* f, const void* x, const void* xs, const void* mask, void* y, int64_t rows,
            int filter, double scale, double* sums, double* partial, hipStream_t st) {
  FilterIO<T> io;
  io.x = (const T*)x;
  io.xs = (const T*)xs;
  io.mask = (const T*)mask;
  io.y = (T*)y;
  io.spec = (const T*)f->spec;
  io.tw = (const typename Cplx<T>::type*)f->tw;
  io.rows = rows;
  io.n = f->n;
  io.scale = (T)scale;
  io.partial = sums ? partial : nullptr;
  const int blocks = apply_blocks(f, rows, filter);
  if (rows > 0) {
    const int pid = prof_begin(NW_K_FILTER, st, rows);
    if (filter)
      dispatch_fft<T>(f->log2L, io, blocks, st);
    else
      hipLaunchKernelGGL((k_filter_ident<T>), dim3(blocks), dim3(kFilterThreads), 0, st, io);
    prof_end(pid, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(NW_ERR_HIP, std::string("k_filter: ") + hipGetErrorString(e));
  }
  if (sums) {  // (no partials when rows == 0: the sums are zero)
    hipLaunchKernelGGL(k_filter_sums, dim3(1), dim3(kFilterThreads), 0, st, partial, blocks,
                       sums);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return fail(NW_ERR_HIP, std::string("k_filter_sums: ") + hipGetErrorString(e));
  }
  return NW_OK;
}

}  // namespace
}  // namespace nw

using nw::fail;

extern "C" {

int32_t nw_filter_create(int32_t dtype, int64_t n, const double* taps, nw_filter** out) {
  if (!out || !taps) return fail(NW_ERR_ARG, "nw_filter_create: null argument");
  *out = nullptr;
  if (dtype != NW_F32 && dtype != NW_F64) return fail(NW_ERR_ARG, "nw_filter_create: bad dtype");
  const int64_t max_n = dtype == NW_F32 ? NW_FILTER_MAX_N_F32 : NW_FILTER_MAX_N_F64;
  if (n < 1 || n > max_n)
    return fail(NW_ERR_UNSUPPORTED, "nw_filter_create: trace length must be in [1, " +
                                        std::to_string(max_n) + "]");
  int log2L = nw::kMinLog2;
  while (((int64_t)1 << log2L) < 2 * n - 1) ++log2L;
  const int64_t L = (int64_t)1 << log2L;
  // the wrapped even kernel and its spectrum (real up to rounding)
  std::vector<std::complex<double>> g((size_t)L, 0.0);
  g[0] = taps[0];
  for (int64_t k = 1; k < n; ++k) {
    g[(size_t)k] = taps[k];
    g[(size_t)(L - k)] = taps[k];
  }
  nw::host_fft(g);
  std::vector<double> spec((size_t)L), tw((size_t)L);
  for (int64_t p = 0; p < L; ++p) spec[(size_t)p] = g[nw::bitrev((size_t)p, log2L)].real() / (double)L;
  for (int64_t k = 0; k < L / 2; ++k) {
    const double ang = -2.0 * nw::kPi * (double)k / (double)L;
    tw[(size_t)(2 * k)] = std::cos(ang);
    tw[(size_t)(2 * k + 1)] = std::sin(ang);
  }
  nw_filter* f = new nw_filter();
  f->dtype = dtype;
  f->n = n;
  f->L = L;
  f->log2L = log2L;
  f->spec = nullptr;
  f->tw = nullptr;
  f->partial = nullptr;
  const int rc = dtype == NW_F32 ? nw::upload<float>(f, spec, tw) : nw::upload<double>(f, spec, tw);
  if (rc != NW_OK) {
    nw_filter_destroy(f);
    return rc;
  }
  *out = f;
  return NW_OK;
}

int32_t nw_filter_destroy(nw_filter* f) {
  if (!f) return NW_OK;
  for (void* p : {f->spec, f->tw, (void*)f->partial})
    if (p) (void)hipFree(p);
  delete f;
  return NW_OK;
}

int64_t nw_filter_length(const nw_filter* f) { return f ? f->L : 0; }

static int check_apply(const char* who, const nw_filter* f, const void* x, int64_t rows) {
  if (!f || (!x && rows > 0)) return fail(NW_ERR_ARG, std::string(who) + ": null argument");
  if (rows < 0) return fail(NW_ERR_ARG, std::string(who) + ": rows must be >= 0");
  if (rows * f->n >= ((int64_t)1 << 40)) return fail(NW_ERR_ARG, std::string(who) + ": too many samples");
  return NW_OK;
}

int32_t nw_filter_apply(nw_filter* f, const void* x, const void* x_sub, const void* mask,
                        void* y, int64_t rows, int32_t filter, double scale, double* sums,
                        void* stream) {
  const int rc = check_apply("nw_filter_apply", f, x, rows);
  if (rc != NW_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (f->dtype == NW_F32)
    return nw::apply_t<float>(f, x, x_sub, mask, y, rows, filter, scale, sums, f->partial, st);
  return nw::apply_t<double>(f, x, x_sub, mask, y, rows, filter, scale, sums, f->partial, st);
}

int32_t nw_filter_apply_noalloc(nw_filter* f, const void* x, const void* x_sub,
                                const void* mask, void* y, int64_t rows, int32_t filter,
                                double scale, double* sums, void* stream) {
  int rc = check_apply("nw_filter_apply_noalloc", f, x, rows);
  if (rc != NW_OK) return rc;
  const int blocks = nw::apply_blocks(f, rows, filter);
  if (2 * (int64_t)blocks > NW_FILTER_NOALLOC_PARTIALS)
    return fail(NW_ERR_ARG, "nw_filter_apply_noalloc: " + std::to_string(blocks) +
                                " workgroups need more partial sums than the " +
                                std::to_string(NW_FILTER_NOALLOC_PARTIALS) + " the array holds");
  double* partial = nullptr;
  if (sums) {
    rc = nw::scratch_on_current_device(&partial);
    if (rc != NW_OK) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  if (f->dtype == NW_F32)
    return nw::apply_t<float>(f, x, x_sub, mask, y, rows, filter, scale, sums, partial, st);
  return nw::apply_t<double>(f, x, x_sub, mask, y, rows, filter, scale, sums, partial, st);
}

}  // extern "C"
