/*
 * leapfrog64.c — FP64 CPU oracle of nefi's acoustic leapfrog (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the checker of libnefi_wave, never part of it: only tests/, the smoke check and
 * the CPU-baseline leg of bench.py load it. It restates the reference step literally, one
 * array operation at a time, in the order of nefi/physics/wave.py:529-552:
 *
 *   lap  = conv(p, K)                                   (wave.py:532, kernel wave.py:125-141)
 *   PML, per axis a:  fd = face_diff(p)                 (wave.py:336-339)
 *                     psi' = alpha*psi + beta*fd        (wave.py:539)
 *                     lap += diff(psi' + psi) * 0.5/h_a (wave.py:540)
 *   rhs  = 2p - b*p_prev + m*lap                        (wave.py:542; b = 1 when absent)
 *   rhs -= zz*p                                         (wave.py:543-544, 2-D PML)
 *   rhs += src_s * w_n                                  (wave.py:545-548, 720-721)
 *   p_next = a*rhs                                      (wave.py:549; a = 1 when absent)
 *   Cerjan: state' = (taper*p, taper*p_next)            (wave.py:550-551)
 *   observations: sum_k w[r,k] p^n[idx[r,k]] at n = j*substeps, j < n_t  (wave.py:321-324,
 *                 726-729, record set wave.py:427-428, n = 0 included: timestep.py:140)
 *
 * with full padded-grid coefficient arrays (a, b, zz, taper, alpha, beta) built by the
 * caller exactly as wave.py:459-504 builds them, and a dense source map per shot
 * (wave.py:312-318, 670-673). The gradient is reverse-mode differentiation of that op
 * sequence by hand, op by op in reverse order, over the stored forward states; it is written
 * independently of the fused reverse kernel it is used to check.
 *
 * Build: oracle/build.sh (gcc, OpenMP over grid rows and cells; every parallel loop writes
 * disjoint elements and sums nothing across them, so results do not depend on the thread
 * count).
 */
#include <stdlib.h>
#include <string.h>

typedef struct {
  int ndim, order, absorbing; /* absorbing: 0 none, 1 sponge, 2 cerjan, 3 pml */
  long dims[3];               /* padded grid, axis 2 fastest, leading unused axes = 1 */
  double h[3];
  int n_steps, n_t, substeps, n_shots;
  const double* a;       /* (cells) or NULL */
  const double* b;       /* (cells) or NULL */
  const double* zz;      /* (cells) or NULL */
  const double* taper;   /* (cells) or NULL */
  const double* alpha[3];/* PML face grids (faces of axis a) or NULL */
  const double* beta[3];
  const double* src;     /* (n_shots, cells), already delta/vol*dt^2, or NULL */
  const double* wavelet; /* (n_steps) */
  int rec_taps, n_rec;   /* receivers shared by every shot */
  const long* rec_idx;   /* (n_rec, rec_taps) */
  const double* rec_w;
} or_problem;

static long cells_of(const or_problem* P) { return P->dims[0] * P->dims[1] * P->dims[2]; }

static int used(const or_problem* P, int a) { return a >= 3 - P->ndim; }

static long fcells_of(const or_problem* P, int a) {
  long n = 1;
  for (int k = 0; k < 3; ++k) n *= P->dims[k] + (k == a);
  return n;
}

/* 1-D coefficients of the cross kernel: centre, +-1, +-2 (wave.py:121) */
static void coeffs(int order, double c[3]) {
  if (order == 4) {
    c[0] = -5.0 / 2.0;
    c[1] = 4.0 / 3.0;
    c[2] = -1.0 / 12.0;
  } else {
    c[0] = -2.0;
    c[1] = 1.0;
    c[2] = 0.0;
  }
}

/* out = K * p (zero values outside the grid); K is symmetric so this is also its transpose.
 * Per cell: out = centre * p + acc, where centre = sum over used axes a of c0/h_a^2 and acc
 * adds, axis by axis (0, 1, 2) and radius by radius (1, 2), (c_k/h_a^2) * (p[-k] + p[+k]) with
 * neighbours outside the grid left out. */
void or_laplacian(const or_problem* P, const double* p, double* out) {
  const long D0 = P->dims[0], D1 = P->dims[1], D2 = P->dims[2];
  double c[3];
  coeffs(P->order, c);
  const int r = P->order == 4 ? 2 : 1;
  double centre = 0.0, ck[3][3] = {{0}};
  for (int a = 0; a < 3; ++a) {
    if (!used(P, a)) continue;
    const double ih2 = 1.0 / (P->h[a] * P->h[a]);
    centre += c[0] * ih2;
    for (int k = 1; k <= r; ++k) ck[a][k] = c[k] * ih2;
  }
  const long st[3] = {D1 * D2, D2, 1};
  const long dims[3] = {D0, D1, D2};
  const int a0 = 3 - P->ndim; /* first used axis */
#pragma omp parallel for collapse(2) schedule(static)
  for (long z = 0; z < D0; ++z)
    for (long y = 0; y < D1; ++y) {
      const long co_zy[2] = {z, y};
      for (long x = 0; x < D2; ++x) {
        const long i = (z * D1 + y) * D2 + x;
        double acc = 0.0;
        for (int a = a0; a < 3; ++a) {
          const long co = a == 2 ? x : co_zy[a];
          for (int k = 1; k <= r; ++k) {
            double s = 0.0;
            if (co - k >= 0) s += p[i - k * st[a]];
            if (co + k < dims[a]) s += p[i + k * st[a]];
            acc += ck[a][k] * s;
          }
        }
        out[i] = centre * p[i] + acc;
      }
    }
}

/* fd_f = p[f] - p[f-1] along axis a on the dims[a]+1 faces, zero outside */
static void face_diff(const or_problem* P, int a, const double* p, double* fd) {
  const long E0 = P->dims[0] + (a == 0), E1 = P->dims[1] + (a == 1), E2 = P->dims[2] + (a == 2);
  const long D1 = P->dims[1], D2 = P->dims[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (long z = 0; z < E0; ++z)
    for (long y = 0; y < E1; ++y)
      for (long x = 0; x < E2; ++x) {
        const long f[3] = {z, y, x};
        long lo[3] = {z, y, x};
        lo[a] -= 1;
        const int hi_in = f[a] < P->dims[a];
        const int lo_in = lo[a] >= 0;
        const double vh = hi_in ? p[(f[0] * D1 + f[1]) * D2 + f[2]] : 0.0;
        const double vl = lo_in ? p[(lo[0] * D1 + lo[1]) * D2 + lo[2]] : 0.0;
        fd[(z * E1 + y) * E2 + x] = vh - vl;
      }
}

/* out_i = q[i+1] - q[i] along axis a (faces -> cells) */
static void diff_faces(const or_problem* P, int a, const double* q, double* out) {
  const long D0 = P->dims[0], D1 = P->dims[1], D2 = P->dims[2];
  const long E1 = D1 + (a == 1), E2 = D2 + (a == 2);
#pragma omp parallel for collapse(2) schedule(static)
  for (long z = 0; z < D0; ++z)
    for (long y = 0; y < D1; ++y)
      for (long x = 0; x < D2; ++x) {
        long f[3] = {z, y, x};
        const long lo = (f[0] * E1 + f[1]) * E2 + f[2];
        f[a] += 1;
        const long hi = (f[0] * E1 + f[1]) * E2 + f[2];
        out[(z * D1 + y) * D2 + x] = q[hi] - q[lo];
      }
}

/* transposes of the two face operators */
static void face_diff_T(const or_problem* P, int a, const double* g, double* out) {
  /* fd_f = p_f - p_{f-1}  =>  out_i = g_i - g_{i+1} */
  const long D0 = P->dims[0], D1 = P->dims[1], D2 = P->dims[2];
  const long E1 = D1 + (a == 1), E2 = D2 + (a == 2);
#pragma omp parallel for collapse(2) schedule(static)
  for (long z = 0; z < D0; ++z)
    for (long y = 0; y < D1; ++y)
      for (long x = 0; x < D2; ++x) {
        long f[3] = {z, y, x};
        const long lo = (f[0] * E1 + f[1]) * E2 + f[2];
        f[a] += 1;
        const long hi = (f[0] * E1 + f[1]) * E2 + f[2];
        out[(z * D1 + y) * D2 + x] = g[lo] - g[hi];
      }
}

static void diff_faces_T(const or_problem* P, int a, const double* g, double* out) {
  /* out_i = q_{i+1} - q_i  =>  qbar_f = g_{f-1} - g_f (zero outside) */
  const long E0 = P->dims[0] + (a == 0), E1 = P->dims[1] + (a == 1), E2 = P->dims[2] + (a == 2);
  const long D1 = P->dims[1], D2 = P->dims[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (long z = 0; z < E0; ++z)
    for (long y = 0; y < E1; ++y)
      for (long x = 0; x < E2; ++x) {
        const long f[3] = {z, y, x};
        long lo[3] = {z, y, x};
        lo[a] -= 1;
        const double gh = f[a] < P->dims[a] ? g[(f[0] * D1 + f[1]) * D2 + f[2]] : 0.0;
        const double gl = lo[a] >= 0 ? g[(lo[0] * D1 + lo[1]) * D2 + lo[2]] : 0.0;
        out[(z * E1 + y) * E2 + x] = gl - gh;
      }
}

static int recorded(const or_problem* P, int n) {
  return n % P->substeps == 0 && n / P->substeps < P->n_t;
}

static void observe(const or_problem* P, const double* p, int s, int j, double* traces) {
  for (int r = 0; r < P->n_rec; ++r) {
    double v = 0.0;
    for (int k = 0; k < P->rec_taps; ++k)
      v += P->rec_w[r * P->rec_taps + k] * p[P->rec_idx[r * P->rec_taps + k]];
    traces[((long)s * P->n_rec + r) * P->n_t + j] = v;
  }
}

/* One step of one shot. prev, cur: state; psi: PML fields (axes 0..2, NULL unused).
 * Writes new_prev, new_cur, new_psi (may not alias the inputs). */
static void step(const or_problem* P, const double* m, int s, int n, const double* prev,
                 const double* cur, double* const* psi, double* new_prev, double* new_cur,
                 double* const* new_psi, double* lap, double* tmpf, double* tmpc) {
  const long N = cells_of(P);
  or_laplacian(P, cur, lap);
  if (P->absorbing == 3)
    for (int a = 0; a < 3; ++a) {
      if (!used(P, a)) continue;
      const long F = fcells_of(P, a);
      face_diff(P, a, cur, tmpf);
#pragma omp parallel for schedule(static)
      for (long f = 0; f < F; ++f) {
        new_psi[a][f] = P->alpha[a][f] * psi[a][f] + P->beta[a][f] * tmpf[f];
        tmpf[f] = new_psi[a][f] + psi[a][f];
      }
      diff_faces(P, a, tmpf, tmpc);
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; ++i) lap[i] += tmpc[i] * (0.5 / P->h[a]);
    }
  const double w = (P->src && n < P->n_steps) ? P->wavelet[n] : 0.0;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < N; ++i) {
    double rhs = 2.0 * cur[i] - (P->b ? P->b[i] * prev[i] : prev[i]) + m[i] * lap[i];
    if (P->zz) rhs -= P->zz[i] * cur[i];
    if (P->src) rhs += P->src[(long)s * N + i] * w;
    const double nxt = P->a ? P->a[i] * rhs : rhs;
    if (P->taper) {
      new_prev[i] = P->taper[i] * cur[i];
      new_cur[i] = P->taper[i] * nxt;
    } else {
      new_prev[i] = cur[i];
      new_cur[i] = nxt;
    }
  }
}

/* Forward run of every shot from (p_prev0, p0) (each (n_shots, cells)); traces (S, R, n_t);
 * final state written to state_out (2, S, cells) if not NULL. */
int or_forward(const or_problem* P, const double* m, const double* p_prev0, const double* p0,
               double* traces, double* state_out) {
  const long N = cells_of(P);
  long Fmax = 0;
  for (int a = 0; a < 3; ++a)
    if (fcells_of(P, a) > Fmax) Fmax = fcells_of(P, a);
  double* buf = (double*)calloc((size_t)(4 * N + 6 * Fmax + Fmax + N + N), sizeof(double));
  if (!buf) return 1;
  double *pa = buf, *pb = buf + N, *qa = buf + 2 * N, *qb = buf + 3 * N;
  double* psi_a[3] = {0, 0, 0};
  double* psi_b[3] = {0, 0, 0};
  double* fb = buf + 4 * N;
  for (int a = 0; a < 3; ++a) {
    psi_a[a] = fb + 2 * a * Fmax;
    psi_b[a] = fb + (2 * a + 1) * Fmax;
  }
  double* tmpf = buf + 4 * N + 6 * Fmax;
  double* lap = tmpf + Fmax;
  double* tmpc = lap + N;
  for (int s = 0; s < P->n_shots; ++s) {
    memcpy(pa, p_prev0 + (long)s * N, N * sizeof(double));
    memcpy(pb, p0 + (long)s * N, N * sizeof(double));
    for (int a = 0; a < 3; ++a) memset(psi_a[a], 0, Fmax * sizeof(double));
    double *prev = pa, *cur = pb, *nprev = qa, *ncur = qb;
    double** psi = psi_a;
    double** npsi = psi_b;
    if (recorded(P, 0)) observe(P, cur, s, 0, traces);
    for (int n = 0; n < P->n_steps; ++n) {
      step(P, m, s, n, prev, cur, psi, nprev, ncur, npsi, lap, tmpf, tmpc);
      double* t;
      t = prev; prev = nprev; nprev = t;
      t = cur; cur = ncur; ncur = t;
      double** u = psi; psi = npsi; npsi = u;
      if (recorded(P, n + 1)) observe(P, cur, s, (n + 1) / P->substeps, traces);
    }
    if (state_out) {
      memcpy(state_out + (long)s * N, prev, N * sizeof(double));
      memcpy(state_out + ((long)P->n_shots + s) * N, cur, N * sizeof(double));
    }
  }
  free(buf);
  return 0;
}

/* Forward run of one shot from (p_prev0, p0) in which, after every step n -> n+1, the cells
 * cells[0..n_cells-1] of p^{n+1} are set to values[(n+1) * n_cells + k]: the receiver clamp of
 * time_reversal's Dirichlet mode (nefi/physics/wave.py:958-962). p0 is taken as given (the
 * caller clamps it for n = 0). The final state is written to state_out (2, cells). */
int or_forward_clamped(const or_problem* P, const double* m, const double* p_prev0,
                       const double* p0, const long* cells, int n_cells, const double* values,
                       double* state_out) {
  const long N = cells_of(P);
  long Fmax = 0;
  for (int a = 0; a < 3; ++a)
    if (fcells_of(P, a) > Fmax) Fmax = fcells_of(P, a);
  double* buf = (double*)calloc((size_t)(4 * N + 6 * Fmax + Fmax + N + N), sizeof(double));
  if (!buf) return 1;
  double *pa = buf, *pb = buf + N, *qa = buf + 2 * N, *qb = buf + 3 * N;
  double* psi_a[3];
  double* psi_b[3];
  double* fb = buf + 4 * N;
  for (int a = 0; a < 3; ++a) {
    psi_a[a] = fb + 2 * a * Fmax;
    psi_b[a] = fb + (2 * a + 1) * Fmax;
  }
  double* tmpf = buf + 4 * N + 6 * Fmax;
  double* lap = tmpf + Fmax;
  double* tmpc = lap + N;
  memcpy(pa, p_prev0, N * sizeof(double));
  memcpy(pb, p0, N * sizeof(double));
  double *prev = pa, *cur = pb, *nprev = qa, *ncur = qb;
  double** psi = psi_a;
  double** npsi = psi_b;
  for (int n = 0; n < P->n_steps; ++n) {
    step(P, m, 0, n, prev, cur, psi, nprev, ncur, npsi, lap, tmpf, tmpc);
    for (int k = 0; k < n_cells; ++k) ncur[cells[k]] = values[(long)(n + 1) * n_cells + k];
    double* t;
    t = prev; prev = nprev; nprev = t;
    t = cur; cur = ncur; ncur = t;
    double** u = psi; psi = npsi; npsi = u;
  }
  memcpy(state_out, prev, N * sizeof(double));
  memcpy(state_out + N, cur, N * sizeof(double));
  free(buf);
  return 0;
}

/* Gradient of J = sum(residual * traces) with respect to m (summed over shots, (cells)) and
 * to the initial state of every shot (grad_prev0, grad_p0: (S, cells); either may be NULL).
 *
 * The reverse of step() is run op by op over stored forward states. With every >= n_steps all
 * states are stored at once; otherwise the forward states at steps 0, every, 2*every, ... are
 * kept and each block of `every` steps is recomputed from its snapshot before its reverse
 * steps (the scheme of torch.utils.checkpoint in nefi/physics/timestep.py:151-169). The
 * recomputation repeats the same operations on the same inputs, so both give identical bits.
 * The reverse step reads the Laplacian (PML face terms included) that the recomputed forward
 * step formed, instead of forming it again. If `traces` is not NULL it receives the traces
 * (S, R, n_t), observed on the recomputed states.
 */
int or_gradient_ckpt(const or_problem* P, const double* m, const double* p_prev0,
                     const double* p0, const double* residual, double* grad_m,
                     double* grad_prev0, double* grad_p0, int every, double* traces) {
  const long N = cells_of(P);
  const int T = P->n_steps;
  const int pml = P->absorbing == 3;
  if (every <= 0 || every > T) every = T > 0 ? T : 1;
  const int nb = T > 0 ? (T + every - 1) / every : 1;
  long F[3] = {0, 0, 0}, Ftot = 0;
  for (int a = 0; a < 3; ++a)
    if (pml && used(P, a)) {
      F[a] = fcells_of(P, a);
      Ftot += F[a];
    }
  const long slot = 2 * N + Ftot; /* prev, cur, psi */
  /* hist: the states of one block; snap: the state at the start of every block */
  double* hist = (double*)malloc((size_t)(every + 1) * slot * sizeof(double));
  double* snap = (double*)malloc((size_t)nb * slot * sizeof(double));
  double* laps = (double*)malloc((size_t)every * N * sizeof(double)); /* lap of each step */
  long Fmax = 1;
  for (int a = 0; a < 3; ++a)
    if (fcells_of(P, a) > Fmax) Fmax = fcells_of(P, a);
  double* work = (double*)calloc((size_t)(10 * N + 8 * Fmax), sizeof(double));
  if (!hist || !snap || !laps || !work) {
    free(hist);
    free(snap);
    free(laps);
    free(work);
    return 1;
  }
  double* lap = work;
  double* tmpc = work + N;
  double* gprev = work + 2 * N;  /* adjoint of state slot 0 */
  double* gcur = work + 3 * N;   /* adjoint of state slot 1 */
  double* nprev_g = work + 4 * N;
  double* ncur_g = work + 5 * N;
  double* lapbar = work + 6 * N;
  double* rhsbar = work + 7 * N;
  double* tmpc2 = work + 8 * N;
  double* tmpf = work + 10 * N;
  double* tmpf2 = tmpf + Fmax;
  double* gpsi[3];
  double* ngpsi[3];
  for (int a = 0; a < 3; ++a) {
    gpsi[a] = tmpf + (2 + 2 * a) * Fmax;
    ngpsi[a] = tmpf + (3 + 2 * a) * Fmax;
  }
  memset(grad_m, 0, N * sizeof(double));
  /* state slots hold (prev, cur, psi of the used axes in order) */
#define SLOT_PSI(base_, psi_)                                    \
  do {                                                           \
    long off_ = 2 * N;                                           \
    for (int a_ = 0; a_ < 3; ++a_) {                             \
      (psi_)[a_] = 0;                                            \
      if (F[a_]) {                                               \
        (psi_)[a_] = (base_) + off_;                             \
        off_ += F[a_];                                           \
      }                                                          \
    }                                                            \
  } while (0)
  for (int s = 0; s < P->n_shots; ++s) {
    /* forward sweep to the start of the last block, keeping every block's first state */
    memcpy(snap, p_prev0 + (long)s * N, N * sizeof(double));
    memcpy(snap + N, p0 + (long)s * N, N * sizeof(double));
    memset(snap + 2 * N, 0, Ftot * sizeof(double));
    for (int b = 1; b < nb; ++b) {
      memcpy(hist, snap + (long)(b - 1) * slot, slot * sizeof(double));
      for (int n = (b - 1) * every; n < b * every; ++n) {
        double* hn = hist + (long)((n - (b - 1) * every) & 1) * slot;
        double* hx = hist + (long)(((n - (b - 1) * every) + 1) & 1) * slot;
        double* psi[3];
        double* npsi[3];
        SLOT_PSI(hn, psi);
        SLOT_PSI(hx, npsi);
        step(P, m, s, n, hn, hn + N, psi, hx, hx + N, npsi, lap, tmpf, tmpc);
      }
      memcpy(snap + (long)b * slot, hist + (long)(every & 1) * slot, slot * sizeof(double));
    }
    /* reverse sweep, block by block */
    memset(gprev, 0, N * sizeof(double));
    memset(gcur, 0, N * sizeof(double));
    for (int a = 0; a < 3; ++a)
      if (F[a]) memset(gpsi[a], 0, F[a] * sizeof(double));
    const double* res = residual + (long)s * P->n_rec * P->n_t;
#define INJECT(n_, g_)                                                               \
  if (recorded(P, n_)) {                                                            \
    const int j_ = (n_) / P->substeps;                                              \
    for (int r = 0; r < P->n_rec; ++r)                                              \
      for (int k = 0; k < P->rec_taps; ++k)                                         \
        (g_)[P->rec_idx[r * P->rec_taps + k]] += P->rec_w[r * P->rec_taps + k] * res[r * P->n_t + j_]; \
  }
    INJECT(T, gcur);
    for (int b = nb - 1; b >= 0; --b) {
      const int n0 = b * every, n1 = (b + 1) * every < T ? (b + 1) * every : T;
      /* recompute the block's states: hist[k] = state n0 + k */
      memcpy(hist, snap + (long)b * slot, slot * sizeof(double));
      if (traces && n0 == 0 && recorded(P, 0)) observe(P, hist + N, s, 0, traces);
      for (int n = n0; n < n1; ++n) {
        double* hn = hist + (long)(n - n0) * slot;
        double* hx = hist + (long)(n - n0 + 1) * slot;
        double* psi[3];
        double* npsi[3];
        SLOT_PSI(hn, psi);
        SLOT_PSI(hx, npsi);
        step(P, m, s, n, hn, hn + N, psi, hx, hx + N, npsi, laps + (long)(n - n0) * N, tmpf,
             tmpc);
        if (traces && recorded(P, n + 1)) observe(P, hx + N, s, (n + 1) / P->substeps, traces);
      }
      for (int n = n1 - 1; n >= n0; --n) {
        /* lap^n of the forward step n, as the recomputation formed it */
        const double* lapn = laps + (long)(n - n0) * N;
        /* outputs: (new_prev, new_cur[, new_psi]) with adjoints (gprev, gcur[, gpsi]) */
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; ++i) {
          double gn = gcur[i];  /* adjoint of nxt (before taper) */
          double gc_direct = gprev[i];
          if (P->taper) {
            gn = P->taper[i] * gcur[i];
            gc_direct = P->taper[i] * gprev[i];
          }
          const double gr = P->a ? P->a[i] * gn : gn; /* rhs adjoint */
          rhsbar[i] = gr;
          /* rhs = 2 cur - b prev + m lap - zz cur + src w */
          ncur_g[i] = gc_direct + 2.0 * gr - (P->zz ? P->zz[i] * gr : 0.0);
          nprev_g[i] = -(P->b ? P->b[i] * gr : gr);
          grad_m[i] += gr * lapn[i];
          lapbar[i] = m[i] * gr;
        }
        /* lap = K cur + sum_a 0.5/h diff(npsi + psi) */
        or_laplacian(P, lapbar, tmpc2);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < N; ++i) ncur_g[i] += tmpc2[i];
        if (pml)
          for (int a = 0; a < 3; ++a) {
            if (!F[a]) continue;
            /* sbar = diff^T(lapbar) * 0.5/h feeds both npsi and psi */
            diff_faces_T(P, a, lapbar, tmpf);
#pragma omp parallel for schedule(static)
            for (long f = 0; f < F[a]; ++f) {
              tmpf[f] *= 0.5 / P->h[a];
              /* total adjoint of npsi: from later steps (gpsi) + this term */
              tmpf2[f] = gpsi[a][f] + tmpf[f];
              /* npsi = alpha psi + beta fd(cur) */
              ngpsi[a][f] = tmpf[f] + P->alpha[a][f] * tmpf2[f];
              tmpf[f] = P->beta[a][f] * tmpf2[f];
            }
            face_diff_T(P, a, tmpf, tmpc);
#pragma omp parallel for schedule(static)
            for (long i = 0; i < N; ++i) ncur_g[i] += tmpc[i];
          }
        double* t;
        t = gprev; gprev = nprev_g; nprev_g = t;
        t = gcur; gcur = ncur_g; ncur_g = t;
        for (int a = 0; a < 3; ++a) {
          t = gpsi[a]; gpsi[a] = ngpsi[a]; ngpsi[a] = t;
        }
        INJECT(n, gcur);
      }
    }
#undef INJECT
    if (grad_prev0) memcpy(grad_prev0 + (long)s * N, gprev, N * sizeof(double));
    if (grad_p0) memcpy(grad_p0 + (long)s * N, gcur, N * sizeof(double));
  }
#undef SLOT_PSI
  free(hist);
  free(snap);
  free(laps);
  free(work);
  return 0;
}

/* Store-all form: every forward state kept, no recomputation. */
int or_gradient(const or_problem* P, const double* m, const double* p_prev0, const double* p0,
                const double* residual, double* grad_m, double* grad_prev0, double* grad_p0) {
  return or_gradient_ckpt(P, m, p_prev0, p0, residual, grad_m, grad_prev0, grad_p0,
                          P->n_steps, NULL);
}
