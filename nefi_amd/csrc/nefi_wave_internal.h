// Kernel parameter blocks shared by the generic and the 2-D kernels of libnefi_wave.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace nw {

extern thread_local std::string g_err;
int fail(int code, const std::string& msg);
// per-launch timing brackets (nw_profile_enable): record index, or -1 when timing is off
int prof_begin(int kind, hipStream_t st, int64_t units = 1);
void prof_end(int id, hipStream_t st);

// Geometry and constants of a plan, passed by value to every kernel.
template <typename T>
struct Geo {
  int64_t D[3];       // padded extents, axis 2 fastest
  int64_t F[3];       // field (unpadded) extents
  int64_t nabs[3];    // absorbing cells per side
  int64_t P2;         // row pitch of axis 2 (>= D[2])
  int64_t cells;      // D0*D1*P2: elements of one shot's wavefield
  int64_t fcells[3];  // cells of the face grid of each axis (rows of axis 2 faces pitched)
  int64_t FP2;        // row pitch of the axis 2 face grid: a multiple of 4, >= D[2] + 1, P2
  int used[3];        // axis carries a stencil
  int nd;             // physical dimensions
  T cc;               // centre coefficient: sum over used axes of c0 / h^2
  T c1[3], c2[3];     // neighbour coefficients per axis (radius 1, 2)
  T inv2h[3];         // 1 / (2 h) per axis (PML divergence)
  T dt2;              // dt^2 in T
  double dt2_d;       // dt^2 in double
  const T* damp[3];   // see nw_desc.damp
  const T* alpha[3];
  const T* kappa[3];
  const T* zf[3];
  const T* zc[3];
  // per axis, the cells [quiet[a][0], quiet[a][1]) whose damping is neutral (sponge 0, Cerjan
  // taper >= 1, every cell without a layer): a tile inside that box steps with a = b = 1.
  // PML: the cells whose damping, cell profile and both face profiles vanish (a tile inside
  // that box has zero face fields)
  int quiet[3][2];
  int damp_kind;  // absorbing layer (nw_absorbing)
};

template <typename T>
struct StepIO {
  const T* m;
  const T* p_prev;
  const T* p;
  T* p_next;
  const T* psi_old[3];
  T* psi_new[3];
  int taper_prev;      // Cerjan: taper p^{n-1} when reading it
  int src_taps;
  const int32_t* src_idx;
  const T* src_val;
  const T* wavelet;
  int n;               // step index (wavelet sample)
  T* ck_prev;          // checkpoint slot (or null)
  T* ck_cur;
  T* ck_psi[3];
  T* lap_out;          // receives the Laplacian of p^n (fast 2-D kernels; or null)
};

template <typename T>
struct SampleIO {
  const T* p;
  T* traces;
  int j, n_t, taps;
  const int32_t* rec_first;
  const int32_t* rec_count;
  const int32_t* trace_first;
  const int32_t* rec_idx;
  const T* rec_w;
};

template <typename T>
struct InjectIO {
  T* lam;
  const T* residual;
  int j, n_t;
  const int32_t* inj_first;
  const int32_t* inj_count;
  const int32_t* inj_cell;
  const int32_t* inj_off;
  const int32_t* inj_rec;
  const T* inj_w;
  const int32_t* trace_first;
};

template <typename T>
struct AdjIO {
  const T* ma;         // m * a_eff
  const T* lam1;       // lam^{n+1}
  const T* lam2;       // lam^{n+2}
  T* lam0;             // lam^n (may alias lam2: read pointwise before the write)
  int taper_next;      // Cerjan: step n+1 tapers its p^{n}
  const T* p_n;        // p^n for the imaging term (or null)
  const T* lap_n;      // or the Laplacian of p^n itself, kept by the forward sweep (fast 2-D)
  const T* psi_low[3];   // psi^{n-1/2} (PML imaging)
  const T* psi_half[3];  // psi^{n+1/2}
  const T* psib_in[3];   // adjoint of psi^{n+1/2}
  T* psib_out[3];        // adjoint of psi^{n-1/2}
  double* grad;        // per-shot accumulators (or null)
};

// Temporally blocked forward steps (2-D float32): levels 1..levels of p^{n0+k}, k = 1..levels,
// computed from p^{n0-1}, p^{n0} in one launch. Level k lands in out[k] when non-null.
constexpr int kMaxDepth = 4;
constexpr int kTapChunks = 8;  // receiver taps a workgroup scans per shot: 256 * kTapChunks

struct BlockIO {
  const float* m;
  const float* p_prev;       // p^{n0-1}
  const float* p;            // p^{n0}
  float* out[kMaxDepth + 1];
  float* lap_out[kMaxDepth + 1];  // level k: the Laplacian of p^{n0+k-1} (or null)
  int levels;                // 1 .. depth of the kernel
  int n0;
  int first_step;            // Cerjan: step n tapers its p^{n-1} when n > first_step
  int n_src;                 // steps n < n_src inject the source
  int src_taps;
  const int32_t* src_idx;
  const float* src_val;
  const float* wavelet;
  float* ck_prev;            // checkpoint slot receiving p^{n0-1}, p^{n0} (or null)
  float* ck_cur;
  // values of p at the receiver taps of levels 1 .. levels-1 (or null):
  // tapbuf[((k-1) * n_shots + s) * tap_stride + r * taps + t]
  float* tapbuf;
  int tap_stride;
  int taps;
  const int32_t* rec_first;
  const int32_t* rec_count;
  const int32_t* rec_idx;
  const int32_t* tap_rows;   // shot s: first (2s) and last (2s + 1) grid row of its taps
};

// Temporally blocked reverse steps (2-D float32): levels k = 1..levels compute lambda^{nh-k}
// from lambda^{nh} (lam1) and lambda^{nh+1} (lam2).
struct AdjBlockIO {
  const float* ma;                  // m * a_eff
  const float* lam1;
  const float* lam2;
  float* out[kMaxDepth + 1];        // level k receives lambda^{nh-k} (or null)
  const float* lap[kMaxDepth + 1];  // level k: Laplacian of p^{nh-k} for the imaging sum (or null)
  double* grad;                     // one float64 accumulator grid per shot group (or null)
  int levels;
  int nh;
  int first_step;                   // Cerjan: step n tapers its p^{n-1} when n > first_step
  // residual injection after level k when inj_rec[k]: injv[((k-1) * n_shots + s) * inj_max + u]
  // is added at cell inj_cell[inj_first[s] + u] (the sums k_inject adds)
  const float* injv;
  int inj_max;
  int inj_rec[kMaxDepth + 1];
  const int32_t* inj_first;
  const int32_t* inj_count;
  const int32_t* inj_cell;
  const int32_t* inj_rows;          // shot s: first (2s) and last (2s + 1) grid row of its cells
};

// fast 2-D float32 kernels (nefi_wave_2d.hip)
bool fast2d_supported(const Geo<float>& g, int absorbing, int dtype);
void fast2d_step(const Geo<float>& g, int order, int absorbing, const StepIO<float>& io, int shots,
                 int ty, int spw, int xcd_map, void* stream);
void fast2d_adjoint(const Geo<float>& g, int order, int absorbing, const AdjIO<float>& io,
                    int shots, int ty, int spw, int xcd_map, void* stream);
// depth 2..kMaxDepth, ty 8/16/32; returns false for a shape it has no kernel for
bool fast2d_block_supported(int ty, int depth);
// Work queue of one plan's blocked forward launches (k4_queue): items (row strip, column
// chunk, first shot, end shot) of queue j in [off[j], off[j + 1]), the queues' counters
// (kXcds + 1 ints, zero between launches) and the workgroups of a launch.
constexpr int kQueues = 8;
struct BlockQueue {
  const int4* items;
  int off[kQueues + 1];
  int* ctr;
  int grid;
};
// queue == nullptr: one workgroup per tile (k4_steps); else the work queue
void fast2d_steps(const Geo<float>& g, int order, int absorbing, const BlockIO& io, int shots,
                  int ty, int depth, int spw, int xcd_map, const BlockQueue* queue, void* stream);
// The queue items of a plan's blocked launches (tiles, depth, shots per workgroup), the last
// split_tail tiles of every queue split into single shots, and the workgroups of the queue
// kernel that fit on one CU.
int fast2d_steps_queue_items(const Geo<float>& g, int order, int absorbing, int ty, int depth,
                             int shots, int spw, int split_tail, std::vector<int4>& items,
                             int* off, int* per_cu);
// traces of levels 1..levels-1 from a BlockIO tap buffer; j[k-1] < 0 skips level k
struct LevelRecords {
  int j[kMaxDepth];
};
bool fast2d_adj_block_supported(int ty, int depth);
void fast2d_adjoint_steps(const Geo<float>& g, int order, int absorbing, const AdjBlockIO& io,
                          int shots, int ty, int depth, int spw, int xcd_map, void* stream);
// fast 3-D float32 kernels (nefi_wave_2d.hip): steps without PML faces (writing the Laplacian
// when io.lap_out is set), reverse steps without an imaging term or with one from kept
// Laplacians (io.lap_n)
bool fast3d_supported(const Geo<float>& g, int absorbing, int dtype);
// fast3d_step with `sample` non-null also gathers the receivers of p^n (io.p) into the traces
void fast3d_step(const Geo<float>& g, int order, int absorbing, const StepIO<float>& io, int shots,
                 int xcd_map, const SampleIO<float>* sample, int rec_max, void* stream);
void fast3d_adjoint(const Geo<float>& g, int order, int absorbing, const AdjIO<float>& io,
                    int shots, int xcd_map, void* stream);
// LDS-tiled 2-D float64 forward step (nefi_wave_2d64.hip)
bool fast2d64_supported(const Geo<double>& g, int absorbing, int dtype);
void fast2d64_step(const Geo<double>& g, int order, int absorbing, const StepIO<double>& io,
                   int shots, int xcd_map, void* stream);
void fast2d_sample_levels(const Geo<float>& g, const SampleIO<float>& io, const float* tapbuf,
                          int tap_stride, int rec_max, int shots, int levels,
                          const LevelRecords& lj, void* stream);

}  // namespace nw
