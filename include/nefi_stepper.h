/* nefi_stepper.h — C ABI of the generic wave stepper (libnefi_wave.so).
 *
 * nefi's bring-your-own-physics integrator `TimeStepper` ships a leapfrog scalar-wave step,
 * `wave_step` (nefi/operators/timestepping.py:167-214): second-order Laplacian with periodic,
 * Neumann (replicated edge), Dirichlet (zero) or absorbing boundaries, the absorbing one being a
 * first-order Mur update of every face (timestepping.py:145-164). These entry points are that
 * step as fused kernels, plus the pieces of its hand-written adjoint.
 *
 * Layout: a field is `batch` consecutive grids of shape[3] (row-major, axis 2 fastest; a
 * d-dimensional grid uses the trailing d axes, unused leading extents are 1). Conventions as in
 * nefi_wave.h: dtype 0 = float, 1 = double for every real buffer; stream-ordered calls on
 * `stream` (hipStream_t as void*); 0 or an error code with nw_last_error() text.
 */
#ifndef NEFI_STEPPER_H
#define NEFI_STEPPER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nw_ts_boundary {
  NW_TS_PERIODIC = 0,
  NW_TS_NEUMANN = 1,
  NW_TS_DIRICHLET = 2,
  NW_TS_ABSORBING = 3 /* Neumann ghosts in the Laplacian, Mur update of the faces */
};

typedef struct nw_ts_grid {
  int32_t ndim;     /* spatial dimensions d (1..3) */
  int32_t dtype;    /* 0 float, 1 double */
  int32_t boundary; /* nw_ts_boundary */
  int32_t reserved;
  int64_t batch;    /* grids per field */
  int64_t shape[3]; /* extents, trailing d used */
  double h[3];      /* spacing of each of the 3 axes (unused axes ignored) */
} nw_ts_grid;

/* out = Laplacian of u with the grid's boundary ghosts (absorbing: Neumann)
 * (replaces `laplacian`, timestepping.py:73-87). The operator is symmetric, so it is also its
 * own adjoint. */
int32_t nw_ts_laplacian(const nw_ts_grid* g, const void* u, void* out, void* stream);

/* u_next = 2u - u_prev + dt^2 (c^2 lap(u) + s), or with damping gamma
 * (2u - (1 - gamma dt/2) u_prev + dt^2 (c^2 lap(u) + s)) / (1 + gamma dt/2), then for the
 * absorbing boundary (if mur != 0) the Mur faces of every axis in order (replaces `wave_step`,
 * timestepping.py:198-214, and `_mur`, :145-164). c has u's shape; s and gamma have u's shape
 * or are NULL. */
int32_t nw_ts_wave_step(const nw_ts_grid* g, double dt, const void* u_prev, const void* u,
                        const void* c, const void* s, const void* gamma, void* u_next,
                        int32_t mur, void* stream);

/* The Mur update of the two faces of one axis (0..2, an axis of the grid with extent >= 3):
 * w[face] = u[next to face] + k (w[next to face] - u[face]), k = (c dt - h)/(c dt + h), c at the
 * face cell (timestepping.py:155-163). */
int32_t nw_ts_mur(const nw_ts_grid* g, int32_t axis, double dt, const void* u, const void* c,
                  void* w, void* stream);

/* Adjoint of nw_ts_mur: given g_w (gradient of the updated w, overwritten with the gradient of
 * w before the update) it adds the u and c contributions to g_u and g_c (g_c may be NULL);
 * w_prev is w before the update (read next to the faces; may be NULL when g_c is NULL). */
int32_t nw_ts_mur_adjoint(const nw_ts_grid* g, int32_t axis, double dt, const void* u,
                          const void* c, const void* w_prev, void* g_w, void* g_u, void* g_c,
                          void* stream);

/* Adjoint of nw_ts_wave_step without the Mur faces, given g_w (gradient of its result):
 * g_u_prev = -(1 - gamma dt/2) g_w / (1 + gamma dt/2) (written), g_u += 2 g + lap(dt^2 c^2 g)
 * with g = g_w / (1 + gamma dt/2), and if not NULL g_c += 2 c dt^2 lap(u) g, g_s = dt^2 g,
 * g_gamma = g dt/2 (u_prev - u_next) (written). */
int32_t nw_ts_wave_step_adjoint(const nw_ts_grid* g, double dt, const void* u_prev,
                                const void* u, const void* c, const void* s, const void* gamma,
                                const void* g_w, void* g_u_prev, void* g_u, void* g_c, void* g_s,
                                void* g_gamma, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NEFI_STEPPER_H */
