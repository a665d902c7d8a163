/*
 * nefi_wave.h — C ABI of the MI355X (gfx950) acoustic leapfrog library `libnefi_wave.so`.
 *
 * The reference path is pure Python: `WaveOperator.forward` / `simulate`
 * (nefi/physics/wave.py:699-750), `WaveInitialConditionOperator.simulate` (wave.py:840-872),
 * the step function `_LeapfrogBase._step_fn` (wave.py:520-554), the receiver gather `_sample`
 * (wave.py:321-324), the time loop `run_timestepping` with block checkpointing
 * (nefi/physics/timestep.py:78-169) and autograd's reverse sweep through that loop.
 * Each entry point below replaces one of those pieces; the comment above it names which.
 *
 * Conventions
 *  - Every buffer is device memory owned by the caller (the Python host allocates it with
 *    PyTorch). The library owns only the plan object.
 *  - Grids are row-major with axis 2 fastest. A d-dimensional padded grid (d = 1..3) is stored
 *    as dims = (1,..,1, n_0, .., n_{d-1}); unused leading axes have extent 1 and no stencil.
 *    Wavefield rows are `pitch` elements apart (see nw_desc.pitch); the sound speed c,
 *    initial pressure p0 and the outputs on the field grid are dense.
 *  - A wavefield buffer holds `n_shots` consecutive grids. PML auxiliary fields are face
 *    grids (extent + 1 along their own axis), also `n_shots` consecutive; rows of the axis 2
 *    face grid are stored at a pitch (a multiple of 4 elements, at least dims[2] + 1 and the
 *    wavefield pitch) so that they start 16-byte aligned, and a face buffer is rounded up to
 *    a multiple of 64 elements, so every slot offset stays 256-byte aligned. nw_face_layout
 *    reports both.
 *  - `dtype` selects float (0) or double (1) for every real-valued buffer of a plan.
 *  - Device buffers (wavefields, face fields, model, traces, slots) must be 16-byte aligned:
 *    the tiled kernels move 16-byte vectors. Allocations from hipMalloc or PyTorch are.
 *  - Every call is stream-ordered on `stream` (a hipStream_t passed as void*; NULL = default
 *    stream) and does not synchronise the host, except `nw_plan_create`.
 *  - Every call returns 0 on success or a nonzero code; `nw_last_error()` returns the message
 *    of the last failure on the calling thread. Nothing aborts the process.
 */
#ifndef NEFI_WAVE_H
#define NEFI_WAVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NW_ABI_VERSION 4

enum nw_status {
  NW_OK = 0,
  NW_ERR_ARG = 1,      /* invalid argument or descriptor */
  NW_ERR_HIP = 2,      /* a HIP runtime call failed */
  NW_ERR_UNSUPPORTED = 3
};

enum nw_absorbing { NW_NONE = 0, NW_SPONGE = 1, NW_CERJAN = 2, NW_PML = 3 };
enum nw_dtype { NW_F32 = 0, NW_F64 = 1 };

/* Plan descriptor: the geometry and constants that `_LeapfrogBase.__init__`/`_static`
 * (wave.py:351-504) and `WaveOperator._extra_static` (wave.py:670-674) derive on the host. */
typedef struct nw_desc {
  int32_t ndim;        /* physical dimensions, 1..3 */
  int32_t dtype;       /* nw_dtype */
  int32_t order;       /* Laplacian accuracy order: 2 or 4 */
  int32_t absorbing;   /* nw_absorbing */
  int64_t dims[3];     /* padded grid extents (leading unused axes = 1) */
  int64_t n_abs[3];    /* absorbing cells per side along each axis (0 for unused axes) */
  int64_t field_dims[3]; /* physical (unpadded) grid = dims - 2*n_abs */
  int64_t pitch;       /* elements between consecutive rows of axis 2 in every wavefield
                        * buffer (>= dims[2]; 0 = dims[2]). A pitch that is a multiple of 32
                        * keeps rows on 128-byte cache-line boundaries. Cell (z, y, x) of a
                        * grid sits at (z*dims[1] + y)*pitch + x; columns x >= dims[2] are
                        * never read or written. Flat cell indices below (src_idx, rec_idx,
                        * inj_cell, nw_clamp's cells) use this pitched numbering. */
  double h[3];         /* grid spacing per axis (1.0 for unused axes) */
  double dt;           /* simulation time step */
  int32_t n_steps;     /* leapfrog steps of one simulation */
  int32_t n_t;         /* trace samples; sample j is taken after j*substeps steps */
  int32_t substeps;
  int32_t n_shots;     /* simulations advanced together (one source each, or none) */

  /* Absorbing-layer profiles (device, `dtype`), one 1-D array per used axis, NULL otherwise.
   * sponge, pml: damp[a][i] = dt/2 * rate_a(i) at cell centres (rate = σ_a or ζ_a);
   * cerjan:      damp[a][i] = exp(-(gamma * xi_a(i))^2).
   * pml only:    pml_alpha[a][f] = (1 - dt ζ_f/2)/(1 + dt ζ_f/2) at the dims[a]+1 faces,
   *              pml_kappa[a][f] = dt / ((1 + dt ζ_f/2) h_a), pml_zf[a][f] = ζ_f,
   *              pml_zc[a][i] = ζ at cell centres. */
  const void* damp[3];
  const void* pml_alpha[3];
  const void* pml_kappa[3];
  const void* pml_zf[3];
  const void* pml_zc[3];

  /* Sources (wave.py:312-318, 670-674, 720-721): per shot `src_taps` flat cell indices into
   * the padded grid and their values delta/cell_volume*dt^2; the injected term at step n is
   * src_val * wavelet[n]. src_taps = 0 means no source (initial-value problems). */
  int32_t src_taps;
  const int32_t* src_idx;  /* (n_shots, src_taps) */
  const void* src_val;     /* (n_shots, src_taps) */
  const void* wavelet;     /* (n_steps) */

  /* Receivers (wave.py:272-300, 321-324): multilinear stencils of rec_taps = 2^ndim taps.
   * Shot s samples the rows rec_first[s] .. rec_first[s]+rec_count[s]-1 of the tables and
   * writes trace rows trace_first[s] ... of the (n_trace_rows, n_t) output. */
  int32_t rec_taps;
  int32_t rec_max;         /* max over shots of rec_count */
  int32_t n_trace_rows;
  const int32_t* rec_first;   /* (n_shots) */
  const int32_t* rec_count;   /* (n_shots) */
  const int32_t* trace_first; /* (n_shots) */
  const int32_t* rec_idx;     /* (rows, rec_taps) flat cell index */
  const void* rec_w;          /* (rows, rec_taps) */

  /* Transpose of the receiver stencils, grouped by distinct cell so the reverse sweep adds
   * residuals without atomics (deterministic): shot s touches cells
   * inj_cell[inj_first[s] .. inj_first[s]+inj_count[s]-1]; entry range
   * inj_off[u] .. inj_off[u+1]-1 of cell u lists (receiver r of the shot, weight). */
  int32_t inj_max;            /* max over shots of inj_count */
  const int32_t* inj_first;   /* (n_shots) */
  const int32_t* inj_count;   /* (n_shots) */
  const int32_t* inj_cell;    /* (U) */
  const int32_t* inj_off;     /* (U+1) */
  const int32_t* inj_rec;     /* (E) */
  const void* inj_w;          /* (E) */
} nw_desc;

typedef struct nw_plan nw_plan;

/* ABI version (NW_ABI_VERSION), sizeof(nw_desc) as compiled into the library (lets a binding
 * check its struct layout), and the last error message of the calling thread. */
int32_t nw_abi_version(void);
int64_t nw_desc_bytes(void);
const char* nw_last_error(void);

/* Validate `desc` and keep a copy of it. Replaces `_LeapfrogBase._static` (wave.py:459-504):
 * the constants stay 1-D profiles and tap lists instead of full padded grids. */
int32_t nw_plan_create(const nw_desc* desc, nw_plan** out);
int32_t nw_plan_destroy(nw_plan* plan);

/* m = pad_replicate(c)^2 * dt^2 on the padded grid (wave.py:718, WaveGrid.pad wave.py:222-240).
 * `c` has field_dims. If `ma` is not NULL it also receives m * a_eff, the model scaled by the
 * damping factor of the step, which the reverse sweep uses. */
int32_t nw_model(nw_plan* plan, const void* c, void* m, void* ma, void* stream);

/* Advance steps n_begin .. n_end-1 of the leapfrog (wave.py:520-554 inside the loop of
 * timestep.py:145-148), fusing the Laplacian, damping, source injection and the receiver
 * gather of wave.py:723-729.
 *
 * state[0], state[1]: p^{n-1}, p^n of every shot on entry; on return the two pointers are
 *   swapped as needed so that they again hold p^{n_end-1}, p^{n_end}.
 * PML face fields inside the layer-free box (cells whose damping, cell profile and both face
 *   profiles vanish along every axis) are identically zero; the fast 2-D kernels neither read
 *   nor write them there. Face buffers handed to nw_forward — psi and its spares, checkpoint
 *   and keep slots — must hold zeros in that box (allocate them zeroed), and the psi_bar
 *   buffers of nw_adjoint finite values.
 * psi: PML only — psi[0..ndim-1] hold the auxiliary face fields psi^{n-1/2} and
 *   psi[ndim..2ndim-1] are spare buffers of the same sizes; the pointer pairs are swapped on
 *   return like `state`. NULL otherwise.
 * traces: (n_trace_rows, n_t) output; sample j = n/substeps is written for every recorded
 *   n in [n_begin, n_end] (p^{n_end} included). NULL skips sampling.
 * ckpt / ckpt_every: if ckpt is not NULL, for every n in [n_begin, n_end) with
 *   n % ckpt_every == 0 the state (p^{n-1}, p^n[, psi]) is written to checkpoint slot
 *   n / ckpt_every (slot size = nw_ckpt_slot_bytes).
 * keep / keep_begin: if keep is not NULL, p^{n} for n in [keep_begin, n_end] are also written
 *   to keep slot n - keep_begin (slot = one wavefield buffer; PML: followed by the psi fields,
 *   except when the slots hold Laplacians, NW_OPT_KEEP_LAP), which the reverse sweep reads
 *   instead of recomputing them. The slot size (nw_keep_slot_bytes) depends on the plan's
 *   tiles and options: query it after changing them.
 * first_step: the step index at which the simulation started (Cerjan tapers p^{n-1} only for
 *   n > first_step). */
int32_t nw_forward(nw_plan* plan, const void* m, void** state, void** psi, int32_t n_begin,
                   int32_t n_end, void* traces, void* ckpt, int32_t ckpt_every, void* keep,
                   int32_t keep_begin, int32_t first_step, void* stream);

/* Tiles of the fast 2-D float32 kernels (no counterpart in the reference; set by the
 * project's autotuner). kind 0 = forward step, 1 = reverse step. rows = rows of the strip one
 * wavefront marches down (4, 8, 16, 32 or 64; 0 selects the generic kernels),
 * shots_per_group = shots one workgroup advances in turn. Plans that the fast kernels cover
 * (2-D, float32, padded width and row pitch multiples of 4) start at rows = 16, 1 shot;
 * others report rows = 0 and refuse nonzero rows. */
int32_t nw_plan_set_tiles(nw_plan* plan, int32_t kind, int32_t rows, int32_t shots_per_group);
int32_t nw_plan_get_tiles(const nw_plan* plan, int32_t kind, int32_t* rows,
                          int32_t* shots_per_group);

/* Plan options (no counterpart in the reference; autotuner knobs).
 * NW_OPT_XCD_MAP: 1 (default) hands every XCD a contiguous run of the fast kernels' tiles so
 * neighbouring row strips share an L2; 0 numbers workgroups linearly.
 * NW_OPT_ROWS_PER_ITER: rows the fast forward step loads before computing (1 default, or 2).
 * NW_OPT_LDS_TILES_STEP / _ADJOINT: 1 runs the LDS-tiled variant of the fast forward / reverse
 *   kernel (whole tile with halo loaded at once; needs rows of 8, 16 or 32); 0 (default) the
 *   register march.
 * NW_OPT_GRAPH: 1 captures every distinct nw_forward / nw_adjoint call (same buffers, step
 *   range and flags) into a hipGraph on a private stream the first time and replays it on the
 *   caller's stream afterwards; 0 (default) launches the kernels one by one (graphs captured
 *   so far are kept for when it is switched back on). Changing tiles or the other options
 *   drops the captured graphs. The whole time loop of a call is one graph launch. A plan
 *   keeps at most 256 graphs: capturing one more first destroys the least recently launched
 *   one, after waiting for its last launch to complete.
 *   While kernel timing (nw_profile_enable) is on, calls run eagerly.
 * NW_OPT_GRAPH_COUNT (read only): graphs the plan currently holds.
 * NW_OPT_TIME_BLOCK: temporal blocking depth T of the fast forward step: 1 (default) one launch
 *   per step; 2..4 advances up to T steps per launch, each LDS tile reading p^{n-1}, p^n and m
 *   once (results bit-identical to T single steps). Used when the forward tiles have rows of
 *   8, 16 or 32 and, for calls that record traces, a shot has at most 2048 receiver taps;
 *   other calls step one by one. Setting T > 1 allocates two wavefields and a small tap
 *   buffer held by the plan.
 * NW_OPT_KEEP_LAP: 1 lets keep buffers hold the Laplacian of p^n instead of p^n whenever both
 *   kernel kinds run the fast 2-D kernels or the tiled 3-D ones (reads back as 1 only then):
 *   the forward kernels write it as they step, and the reverse kernels read it pointwise for
 *   the imaging term (same values, no stencil). A keep sweep feeding such a reverse sweep over steps [n0, n1)
 *   must then run forward over [n0, n1) (not n1 - 1). 0 keeps p^n: a blocked keep sweep then
 *   writes only the levels themselves, which measured faster overall. Default: 1 for PML
 *   plans (the kept Laplacian replaces p^n and both face-field levels in the reverse step) and
 *   for plans on the tiled 3-D kernels (whose reverse step images only from kept Laplacians),
 *   0 otherwise.
 * NW_OPT_TIME_BLOCK_ADJOINT: temporal blocking depth of the fast reverse step (1 default,
 *   2..4): up to that many reverse steps per launch, with the residual injections of recorded
 *   steps and the imaging sum inside the launch (lambda bit-identical to single steps; the
 *   float64 imaging sums are added in another order). Used when the reverse tiles have rows of
 *   8 or 16, the imaging term (if any) reads kept Laplacians (NW_OPT_KEEP_LAP) and a shot has
 *   at most 2048 injection cells; other calls step one by one. Setting T > 1 allocates two
 *   wavefields of its own (not those of NW_OPT_TIME_BLOCK), so a blocked reverse sweep may run
 *   beside a blocked forward sweep of the same plan on another stream.
 * NW_OPT_FAST_3D: 1 (default where they cover the plan: 3-D, float32, sponge/Cerjan/none,
 *   padded width and row pitch multiples of 4) runs the 2.5-D tiled 3-D kernels for forward
 *   steps and for reverse steps without an imaging term; 0 the generic kernels. Plans they do
 *   not cover report 0 and refuse 1.
 * NW_OPT_FAST_F64: 1 (default where it covers the plan: 2-D or 3-D, float64, sponge/Cerjan/
 *   none, padded width and row pitch even) runs the LDS-tiled float64 forward step (2.5-D
 *   tiles in 3-D); 0 the generic one. Plans it does not cover report 0 and refuse 1.
 * NW_OPT_WORK_QUEUE: how temporally blocked forward launches (NW_OPT_TIME_BLOCK > 1) are
 *   scheduled. A work queue runs as many workgroups as the device holds at once, each taking
 *   tiles (and, at the end, single shots of tiles) from per-XCD queues until all are empty;
 *   otherwise there is one workgroup per tile. 1 (default): the queue for calls that fill no
 *   keep buffer (a keep sweep may run beside reverse steps on another stream, which the
 *   queue's resident workgroups would keep off the CUs); 2: for every blocked call; 0: never.
 *   Results are identical. The queue's items are built when the forward tiles, the depth or
 *   this option are set (which waits for the device), and its counters are the plan's: calls
 *   on one plan must not run concurrently (as for its blocking wavefields).
 * NW_OPT_WORK_QUEUE_ITEMS (read only): items of the plan's current work queue; 0 when blocked
 *   launches take one workgroup per tile (queue off, no blocking, or no device). */
enum nw_option {
  NW_OPT_XCD_MAP = 0,
  NW_OPT_ROWS_PER_ITER = 1,
  NW_OPT_LDS_TILES_STEP = 2,
  NW_OPT_LDS_TILES_ADJOINT = 3,
  NW_OPT_GRAPH = 4,
  NW_OPT_GRAPH_COUNT = 5,
  NW_OPT_TIME_BLOCK = 6,
  NW_OPT_KEEP_LAP = 7,
  NW_OPT_TIME_BLOCK_ADJOINT = 8,
  NW_OPT_FAST_3D = 9,
  NW_OPT_FAST_F64 = 10,
  NW_OPT_WORK_QUEUE = 11,
  NW_OPT_WORK_QUEUE_ITEMS = 12
};
int32_t nw_plan_set_option(nw_plan* plan, int32_t option, int32_t value);
int32_t nw_plan_get_option(const nw_plan* plan, int32_t option, int32_t* value);

/* Number of float64 gradient accumulator grids nw_adjoint writes (grad_m holds
 * nw_grad_slices * padded cells doubles; nw_model_grad sums all of them). */
int64_t nw_grad_slices(const nw_plan* plan);

/* Bytes of one checkpoint slot and of one keep slot (the latter for the current tiles and
 * options: slots of kept Laplacians hold one wavefield buffer, PML face fields included in it). */
int64_t nw_ckpt_slot_bytes(const nw_plan* plan);
int64_t nw_keep_slot_bytes(const nw_plan* plan);

/* Layout of the plan's PML face buffer along grid axis `axis` (0..2): *elems = elements of one
 * buffer (all shots, rounded up to a multiple of 64; 0 when the plan has no faces along the
 * axis), *row_pitch = elements between consecutive rows of one shot's face grid (dims[2] for
 * the faces of axes 0 and 1; for axis 2 a multiple of 4, >= dims[2] + 1 and >= the wavefield
 * pitch; the elements past face dims[2] of a row are never read or written). The face grid of
 * axis 2 for shot s starts at s * dims[0] * dims[1] * row_pitch. New in ABI version 4 (which
 * pitched the axis 2 face rows); the reference has no counterpart: nefi keeps its face fields
 * as PyTorch tensors (nefi/physics/wave.py:529-541). */
int32_t nw_face_layout(const nw_plan* plan, int32_t axis, int64_t* elems, int64_t* row_pitch);

/* Reverse sweep over steps n_end-1 down to n_begin: the discrete adjoint of the leapfrog that
 * autograd derives from wave.py:520-554 and replays through timestep.py:151-169.
 *
 * lam[0], lam[1]: adjoints of p^{n_end}, p^{n_end+1} on entry (zero at the end of time);
 *   on return they hold the adjoints of p^{n_begin}, p^{n_begin+1}, pointers swapped as needed.
 * psi_bar: PML only, 2*ndim buffers like `psi` of nw_forward (adjoints of psi^{n+1/2}).
 * ma: m * a_eff from nw_model. m: the model itself.
 * keep / keep_begin: p^n (and psi) for n in [keep_begin, ...] as written by nw_forward.
 * residual: (n_trace_rows, n_t) = dJ/dtraces; injected at every recorded n in
 *   [n_begin, n_end) (the caller injects n_end's sample before the first call via the same
 *   rule: pass inject_end = 1).
 * grad_m: if not NULL, float64 accumulators (nw_grad_slices, padded cells) receive
 *   sum_n a_eff*lam^{n+1} * lap(p^n), one grid per shot group.
 * first_step: as in nw_forward. */
int32_t nw_adjoint(nw_plan* plan, const void* m, const void* ma, void** lam, void** psi_bar,
                   int32_t n_begin, int32_t n_end, const void* keep, int32_t keep_begin,
                   const void* residual, int32_t inject_end, double* grad_m,
                   int32_t first_step, void* stream);

/* dJ/dc = 2 dt^2 c * fold(sum_shots grad_m): the adjoint of m = pad_replicate(c)^2 dt^2, with
 * the absorbing layer folded back onto the edge cells (adjoint of wave.py:222-240). `grad_c`
 * has field_dims and the plan's dtype. */
int32_t nw_model_grad(nw_plan* plan, const void* c, const double* grad_m, void* grad_c,
                      void* stream);

/* Initial state of the initial-value operator (wave.py:849-850): p^0 = pad_zeros(p0),
 * p^{-1} = p^0 + m/2 * lap(p^0). `p0` is (n_shots, field_dims). */
int32_t nw_ic_state(nw_plan* plan, const void* m, const void* p0, void* p_prev, void* p,
                    void* stream);

/* Adjoint of nw_ic_state: grad_p0 = crop(lam0 + pbar + lap(m/2 * pbar)) where lam0, pbar are
 * the adjoints of p^0 and p^{-1} left by nw_adjoint. */
int32_t nw_ic_grad(nw_plan* plan, const void* m, const void* lam0, const void* pbar,
                   void* grad_p0, void* stream);

/* Adjoint of p^{-1} after the last reverse step: pbar = -b * a_eff * lam^1. */
int32_t nw_prev_adjoint(nw_plan* plan, const void* lam1, void* pbar, void* stream);

/* out = lap(p) over `count` grids of the plan's padded shape (pitched rows), zero outside
 * (wave.py:144-168). */
int32_t nw_laplacian(nw_plan* plan, const void* p, void* out, int64_t count, void* stream);

/* Set p[s, cells[k]] = values[k] for k < n_cells in every shot: the receiver clamp of
 * time_reversal's Dirichlet mode (wave.py:958-962). */
int32_t nw_clamp(nw_plan* plan, void* p, const int32_t* cells, const void* values,
                 int32_t n_cells, void* stream);

/* Kernel timing (no counterpart in the reference; measurement support for benchmarks).
 * While enabled, every launch of the step, reverse-step, sample and inject kernels is bracketed
 * by HIP events on its own stream, and calls run eagerly even in graph mode. nw_profile_read
 * waits for the recorded events and returns the summed duration (ms) and launch count of one
 * kernel kind since profiling was enabled: 0 = forward step, 1 = reverse step,
 * 2 = receiver sample, 3 = residual inject, 4 = temporally blocked forward steps,
 * 5 = temporally blocked reverse steps, 6 = trace filter (nefi_filter.h).
 * nw_profile_units returns the time steps those launches advanced (for kind 2, the time
 * samples recorded; for kind 6, the traces filtered). Enabling clears earlier records.
 * Process-wide; not thread-safe. */
enum nw_kernel_kind {
  NW_K_STEP = 0,
  NW_K_ADJOINT = 1,
  NW_K_SAMPLE = 2,
  NW_K_INJECT = 3,
  NW_K_STEPS = 4,
  NW_K_ADJOINT_STEPS = 5,
  NW_K_FILTER = 6
};
int32_t nw_profile_enable(int32_t on);
int32_t nw_profile_read(int32_t kind, double* total_ms, int64_t* launches);
int32_t nw_profile_units(int32_t kind, int64_t* units);

#ifdef __cplusplus
}
#endif

#endif /* NEFI_WAVE_H */
