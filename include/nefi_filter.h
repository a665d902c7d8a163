/* nefi_filter.h — C ABI of the trace filters and the fused trace misfit (libnefi_wave.so).
 *
 * The reference filters traces along time with a zero-padded real FFT and a real, even
 * frequency response H:
 *   y = irfft(rfft(x, N) * H, N)[..., :n]
 * (`lowpass`, nefi/instances/wave_fwi/losses.py:29-50, N = next_fast_len(2n), H a cosine-tapered
 * box; `gaussian_lowpass`, nefi/instances/photoacoustic3d/sensor.py:24-38, N = 2n,
 * H = exp(-(f/f_c)^2/2)). With N >= 2n - 1 that map is the symmetric Toeplitz operator
 *   y[t] = sum_{s<n} h[|t - s|] x[s],   h = irfft(H, N)[:n],
 * which is its own adjoint. A filter plan holds h as the spectrum of its circular extension on
 * a power-of-two length L >= 2n - 1; the kernels apply it with an in-LDS FFT, two traces per
 * complex transform, so the result is the reference's filter to rounding.
 *
 * Conventions as in nefi_wave.h: row-major buffers, dtype 0 = float, 1 = double for every real
 * buffer of a plan, stream-ordered calls on `stream` (hipStream_t as void*), int return codes
 * with nw_last_error() text, nothing aborts.
 */
#ifndef NEFI_FILTER_H
#define NEFI_FILTER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nw_filter nw_filter;

/* Largest trace length of a plan: L = 2^ceil(log2(2n - 1)) complex values must fit in LDS. */
#define NW_FILTER_MAX_N_F32 8192
#define NW_FILTER_MAX_N_F64 4096

/* Create a filter plan for traces of n samples from the impulse response taps h[0..n-1]
 * (host, double): the time-domain kernel of the reference's FFT filter at lags 0..n-1
 * (replaces the per-call window + rfft/irfft of losses.py:29-50 and sensor.py:24-38 with a
 * plan built once). Allocates device memory; synchronous. */
int32_t nw_filter_create(int32_t dtype, int64_t n, const double* taps, nw_filter** out);
int32_t nw_filter_destroy(nw_filter* f);

/* FFT length L of a plan (a power of two >= 2n - 1). */
int64_t nw_filter_length(const nw_filter* f);

/* For every row r < rows of the (rows, n) buffers:
 *   u = x[r] - x_sub[r]            (x_sub NULL: u = x[r])
 *   v = filter(u)                  (filter = 0: v = u, no transform)
 *   y[r] = scale * mask[r] * v     (mask NULL: weight 1; y NULL: nothing is written)
 * and, if sums is not NULL, sums[0] = sum over all rows and samples of mask * v^2 and
 * sums[1] = sum of mask (double, device memory, overwritten; summed in a fixed order, so
 * repeated calls give identical bits). `mask` has the shape of x (the caller expands
 * broadcast masks). y may alias x. This one call is the residual, filter, mask and the two
 * masked means of BandLimitedMSE (losses.py:65-80; masked_mean, nefi/measurement.py:86-91);
 * its adjoint source is a second call on mask * v with the gradient scale. */
int32_t nw_filter_apply(nw_filter* f, const void* x, const void* x_sub, const void* mask,
                        void* y, int64_t rows, int32_t filter, double scale, double* sums,
                        void* stream);

/* Allocation-free form of nw_filter_apply. It has no counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc), which create their plans beforehand. Arguments,
 * results and bits are those of nw_filter_apply; the float64 partial sums of the workgroups go
 * to a fixed array of NW_FILTER_NOALLOC_PARTIALS doubles (32 KiB) that the library holds on each
 * GPU it runs on until the process ends, instead of the plan's own buffer. The call allocates
 * nothing. Every call on one GPU shares that array and leaves the last call's partial sums in
 * it (each call writes every partial it reads, so nothing needs clearing): calls that overlap
 * on one GPU (other streams or threads) race, and the caller must order them. A call that
 * would need more partial sums than the array holds is refused with NW_ERR_ARG. */
#define NW_FILTER_NOALLOC_PARTIALS 4096
int32_t nw_filter_apply_noalloc(nw_filter* f, const void* x, const void* x_sub,
                                const void* mask, void* y, int64_t rows, int32_t filter,
                                double scale, double* sums, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NEFI_FILTER_H */
