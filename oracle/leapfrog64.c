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
 * Build: oracle/build.sh (gcc, optional OpenMP over grid rows).
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

/* out = K * p (zero values outside the grid); K is symmetric so this is also its transpose */
void or_laplacian(const or_problem* P, const double* p, double* out) {
  const long D0 = P->dims[0], D1 = P->dims[1], D2 = P->dims[2];
  double c[3];
  coeffs(P->order, c);
  const int r = P->order == 4 ? 2 : 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (long z = 0; z < D0; ++z)
    for (long y = 0; y < D1; ++y)
      for (long x = 0; x < D2; ++x) {
        const long co[3] = {z, y, x};
        const long i = (z * D1 + y) * D2 + x;
        double centre = 0.0, acc = 0.0;
        for (int a = 0; a < 3; ++a) {
          if (!used(P, a)) continue;
          const double ih2 = 1.0 / (P->h[a] * P->h[a]);
          centre += c[0] * ih2;
          const long st = a == 0 ? D1 * D2 : (a == 1 ? D2 : 1);
          for (int k = 1; k <= r; ++k) {
            double s = 0.0;
            if (co[a] - k >= 0) s += p[i - k * st];
            if (co[a] + k < P->dims[a]) s += p[i + k * st];
            acc += c[k] * ih2 * s;
          }
        }
        out[i] = centre * p[i] + acc;
      }
}

/* fd_f = p[f] - p[f-1] along axis a on the dims[a]+1 faces, zero outside */
static void face_diff(const or_problem* P, int a, const double* p, double* fd) {
  const long E0 = P->dims[0] + (a == 0), E1 = P->dims[1] + (a == 1), E2 = P->dims[2] + (a == 2);
  const long D1 = P->dims[1], D2 = P->dims[2];
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
      for (long f = 0; f < F; ++f) new_psi[a][f] = P->alpha[a][f] * psi[a][f] + P->beta[a][f] * tmpf[f];
      for (long f = 0; f < F; ++f) tmpf[f] = new_psi[a][f] + psi[a][f];
      diff_faces(P, a, tmpf, tmpc);
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

/* Gradient of J = sum(residual * traces) with respect to m (summed over shots, (cells)) and
 * to the initial state of every shot (grad_prev0, grad_p0: (S, cells); either may be NULL).
 * Stores every forward state, then runs the reverse of step() op by op. */
int or_gradient(const or_problem* P, const double* m, const double* p_prev0, const double* p0,
                const double* residual, double* grad_m, double* grad_prev0, double* grad_p0) {
  const long N = cells_of(P);
  const int T = P->n_steps;
  const int pml = P->absorbing == 3;
  long F[3] = {0, 0, 0}, Ftot = 0;
  for (int a = 0; a < 3; ++a)
    if (pml && used(P, a)) {
      F[a] = fcells_of(P, a);
      Ftot += F[a];
    }
  const long slot = 2 * N + Ftot; /* prev, cur, psi */
  double* hist = (double*)malloc((size_t)(T + 1) * slot * sizeof(double));
  long Fmax = 1;
  for (int a = 0; a < 3; ++a)
    if (fcells_of(P, a) > Fmax) Fmax = fcells_of(P, a);
  double* work = (double*)calloc((size_t)(10 * N + 8 * Fmax), sizeof(double));
  if (!hist || !work) {
    free(hist);
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
  for (int s = 0; s < P->n_shots; ++s) {
    /* forward, storing state n in hist[n] */
    double* h0 = hist;
    memcpy(h0, p_prev0 + (long)s * N, N * sizeof(double));
    memcpy(h0 + N, p0 + (long)s * N, N * sizeof(double));
    memset(h0 + 2 * N, 0, Ftot * sizeof(double));
    for (int n = 0; n < T; ++n) {
      double* hn = hist + (long)n * slot;
      double* hx = hist + (long)(n + 1) * slot;
      double* psi[3] = {0, 0, 0};
      double* npsi[3] = {0, 0, 0};
      long off = 2 * N;
      for (int a = 0; a < 3; ++a)
        if (F[a]) {
          psi[a] = hn + off;
          npsi[a] = hx + off;
          off += F[a];
        }
      step(P, m, s, n, hn, hn + N, psi, hx, hx + N, npsi, lap, tmpf, tmpc);
    }
    /* reverse sweep */
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
    for (int n = T - 1; n >= 0; --n) {
      const double* hn = hist + (long)n * slot;
      const double* cur = hn + N;
      const double* psi[3] = {0, 0, 0};
      const double* npsi[3] = {0, 0, 0};
      long off = 2 * N;
      for (int a = 0; a < 3; ++a)
        if (F[a]) {
          psi[a] = hn + off;
          npsi[a] = hist + (long)(n + 1) * slot + off;
          off += F[a];
        }
      /* recompute lap^n of the forward step */
      or_laplacian(P, cur, lap);
      if (pml)
        for (int a = 0; a < 3; ++a) {
          if (!F[a]) continue;
          for (long f = 0; f < F[a]; ++f) tmpf[f] = npsi[a][f] + psi[a][f];
          diff_faces(P, a, tmpf, tmpc);
          for (long i = 0; i < N; ++i) lap[i] += tmpc[i] * (0.5 / P->h[a]);
        }
      /* outputs: (new_prev, new_cur[, new_psi]) with adjoints (gprev, gcur[, gpsi]) */
      memset(nprev_g, 0, N * sizeof(double));
      memset(ncur_g, 0, N * sizeof(double));
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
        grad_m[i] += gr * lap[i];
        lapbar[i] = m[i] * gr;
      }
      /* lap = K cur + sum_a 0.5/h diff(npsi + psi) */
      or_laplacian(P, lapbar, tmpc2);
      for (long i = 0; i < N; ++i) ncur_g[i] += tmpc2[i];
      if (pml)
        for (int a = 0; a < 3; ++a) {
          if (!F[a]) continue;
          /* sbar = diff^T(lapbar) * 0.5/h feeds both npsi and psi */
          diff_faces_T(P, a, lapbar, tmpf);
          for (long f = 0; f < F[a]; ++f) tmpf[f] *= 0.5 / P->h[a];
          /* total adjoint of npsi: from later steps (gpsi) + this term */
          for (long f = 0; f < F[a]; ++f) tmpf2[f] = gpsi[a][f] + tmpf[f];
          /* npsi = alpha psi + beta fd(cur) */
          for (long f = 0; f < F[a]; ++f) ngpsi[a][f] = tmpf[f] + P->alpha[a][f] * tmpf2[f];
          for (long f = 0; f < F[a]; ++f) tmpf[f] = P->beta[a][f] * tmpf2[f];
          face_diff_T(P, a, tmpf, tmpc);
          for (long i = 0; i < N; ++i) ncur_g[i] += tmpc[i];
        }
      memcpy(gprev, nprev_g, N * sizeof(double));
      memcpy(gcur, ncur_g, N * sizeof(double));
      for (int a = 0; a < 3; ++a)
        if (F[a]) memcpy(gpsi[a], ngpsi[a], F[a] * sizeof(double));
      INJECT(n, gcur);
    }
#undef INJECT
    if (grad_prev0) memcpy(grad_prev0 + (long)s * N, gprev, N * sizeof(double));
    if (grad_p0) memcpy(grad_p0 + (long)s * N, gcur, N * sizeof(double));
  }
  free(hist);
  free(work);
  return 0;
}
