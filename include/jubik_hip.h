/* C ABI of the MI355X (gfx950) implementation of J-UBIK's forward model and
 * gradient: the spatially variant PSF (linear patching), exposure, read-out,
 * Poisson likelihood and the correlated-field GP sky, plus a fused gradient
 * evaluation and the MGVI metric product.
 *
 * Conventions
 *   - Every entry point returns 0 on success or a negative status
 *     (JUBIK_EINVAL, JUBIK_EUNSUPPORTED, JUBIK_EHIP, JUBIK_ENOMEM,
 *     JUBIK_ENOTREADY, JUBIK_EDOMAIN); jubik_last_error() gives the message
 *     of the calling thread's last failure.
 *   - Array arguments are device pointers owned by the caller, C-contiguous,
 *     typed by the plan's dtype (0 = float32, 1 = float64) unless stated.
 *   - ``stream`` is a hipStream_t (NULL = the default stream).  Execution
 *     entry points only enqueue work: no host synchronisation, no device
 *     allocation.  Plans own their workspace; a plan must not be used from
 *     two host threads at once.
 *   - Citations name the reference function each entry point replaces
 *     (paths relative to the J-UBIK source tree).
 */
#ifndef JUBIK_HIP_H
#define JUBIK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JUBIK_OK 0
#define JUBIK_EINVAL (-1)
#define JUBIK_EUNSUPPORTED (-2)
#define JUBIK_EHIP (-3)
#define JUBIK_ENOMEM (-4)
#define JUBIK_ENOTREADY (-5)
#define JUBIK_EDOMAIN (-6) /* lambda <= 0 at a read-out pixel */

#define JUBIK_FLOAT32 0
#define JUBIK_FLOAT64 1

#define JUBIK_KERNEL_FULL 0   /* (M, P, [E,] H, W) full-frame kernels */
#define JUBIK_KERNEL_WINDOW 1 /* (M, P, [E,] 2dx, 2dx) cut windows */

/* Bytes of the FP64 workspace jubik_poisson_nll_grad needs. */
#define JUBIK_POISSON_WORK_BYTES 8192

const char* jubik_last_error(void);
int jubik_abi_version(void);

/* Kernel timing (no counterpart in the reference, which has no profiling):
 * while enabled, every kernel launch is bracketed by HIP events on its
 * stream; jubik_timing_report synchronises them and writes one line per
 * kernel, "name count total_ms min_ms max_ms", then forgets them. */
int jubik_timing_enable(int on);
int jubik_timing_report(char* buf, int64_t len);
/* Restrict the timing to the kernels named ``name`` (NULL or "": every
 * kernel), so one kernel can be timed without events between the other
 * launches of an evaluation. */
int jubik_timing_select(const char* name);
/* Fault injection for tests of the error paths (no counterpart in the
 * reference): the nth next device allocation the library makes fails with
 * JUBIK_ENOMEM as if the device were full (nth = 0 disarms).  Creation calls
 * release what they allocated before the failure. */
int jubik_debug_fail_alloc(int64_t nth);

/* ---------------------------------------------------------------- linpatch
 * Replaces jubik/convolve.py:66-145 linpatch_convolve(x, domain, kernel,
 * n_patches_per_axis, margin), including the module vmap of its callers
 * (jubik/instruments/chandra/chandra_response.py:127-130,
 * jubik/instruments/erosita/erosita_response.py:209-212).
 * Grid (H, W), E energies, M modules.  kernel_per_energy: 1 if kernels are
 * (M, P, E, ...), 0 if one kernel per patch serves all energies (the
 * broadcast branch of convolve.py:167-171).  x_shared: 1 if one input
 * (E, H, W) feeds all modules (the vmap in_axes=None of x), 0 if the input
 * is (M, E, H, W).                                                        */
typedef struct jubik_linpatch jubik_linpatch;

int jubik_linpatch_create(jubik_linpatch** plan, int dtype, int64_t H, int64_t W, int64_t E,
                          int64_t M, int64_t n_patches_per_axis, int64_t margin,
                          int kernel_per_energy, int x_shared);
int jubik_linpatch_destroy(jubik_linpatch* plan);
/* out[0..7] = dx, window 2dx, FFT length L, P, slicing pad dx/2, jobs M*E,
 * workspace bytes, spectra bytes. */
int jubik_linpatch_info(const jubik_linpatch* plan, int64_t* out);
/* Kernel preparation of convolve.py:104-131 (fftshift, centred cut, margin
 * pad, ifftshift, unit-sum normalisation) and its FFT; kept in the plan. */
int jubik_linpatch_set_kernel(jubik_linpatch* plan, const void* kernel, int layout,
                              void* stream);
/* y (M, E, H, W) = linpatch_convolve(scale * x). */
int jubik_linpatch_forward(jubik_linpatch* plan, const void* x, void* y, double scale,
                           void* stream);
/* gx = scale * linpatch_convolve^T(gy); the VJP JAX derives for
 * convolve.py:66-145 (sums over modules when x_shared). */
int jubik_linpatch_adjoint(jubik_linpatch* plan, const void* gy, void* gx, double scale,
                           void* stream);
/* Tile autotuning (no counterpart in the reference, whose FFTs XLA plans):
 * the forward and adjoint column passes run V / col_div columns per
 * work-group and the signal and cotangent row passes VR / row_div row pairs,
 * col_div and row_div = 1, 2 or 4 (a divisor is compiled in when V / col_div
 * >= 2, VR / row_div >= 1).  Results agree to rounding across tiles (the
 * same butterflies run per column or row; multiply-adds contract differently
 * per tile).  jubik_linpatch_autotune times one forward + adjoint per column
 * tile (row tile 1, reps repetitions, zero inputs, after the kernel is set),
 * then per row tile with the fastest column tile, keeps the fastest pair in
 * the plan and returns it; ms[0..2] receive the milliseconds of column
 * divisors 1, 2, 4 and ms[3..5] those of row divisors 1, 2, 4 (-1 where not
 * compiled in).  It allocates and synchronises: call it at set-up. */
int jubik_linpatch_set_tile(jubik_linpatch* plan, int col_div, int row_div);
int jubik_linpatch_autotune(jubik_linpatch* plan, int reps, int* col_div, int* row_div,
                            double* ms, void* stream);
/* Column-pass transform (no counterpart in the reference, whose FFTs XLA
 * runs in the input precision): kind 0 = the plan's FFT (default), kind 1
 * = each 384-point column DFT as three 128-point DFTs on the matrix cores
 * (v_mfma_f32_32x32x16_bf16, every product split into BF16 hi and lo
 * halves, hi*hi + hi*lo + lo*hi, FP32 accumulation) and an FP32 radix-3
 * step.  Kind 1 needs a float32 plan with window + 2 * margin = 384
 * (JUBIK_EUNSUPPORTED otherwise) and allocates 128 KiB of DFT matrices on
 * first use.  Row passes and the overlap-add are unchanged.  It applies to
 * every later call of the plan and of the models built on it; graphs
 * captured earlier keep the transform they were captured with. */
int jubik_linpatch_set_transform(jubik_linpatch* plan, int kind);

/* ---------------------------------------------------------------- primitives
 * slice_patches of jubik/convolve.py:28-63 on x (B, H, W): out (n^2, B,
 * 2dx+2m, 2dy+2m), dx = (H - 2 margin) / n, patch p = iy * n + ix, window
 * starts clamped as jax.lax.dynamic_slice does.  The adjoint is its
 * transpose (the jax.linear_transpose used at jubik/convolve.py:138-143). */
int jubik_slice_patches(int dtype, const void* x, int64_t B, int64_t H, int64_t W, int64_t n,
                        int64_t margin, void* out, void* stream);
int jubik_slice_patches_adjoint(int dtype, const void* patches, int64_t B, int64_t H, int64_t W,
                                int64_t n, int64_t margin, void* x, void* stream);
/* convolve of jubik/convolve.py:148-176 over two axes of real H x W
 * arrays (the last two after the caller moves them there): out[b] = dvol *
 * Re(IFFT2(FFT2(x[b / (nby / nbx)]) * FFT2(y[b]))), the circular
 * convolution of the reference's fftn / ifftn.  Any H, W up to 2048: a
 * length without its own FFT plan runs zero-padded at a compiled length
 * >= 2 n - 1 and is folded back (the same values).  ``work`` holds
 * jubik_convolve_work_bytes(dtype, nby, H, W) bytes of device memory (-1:
 * unsupported sizes). */
int64_t jubik_convolve_work_bytes(int dtype, int64_t nby, int64_t H, int64_t W);
int jubik_convolve(int dtype, const void* x, const void* y, int64_t nbx, int64_t nby, int64_t H,
                   int64_t W, double dvol, void* work, void* out, void* stream);

/* ---------------------------------------------------------------- response
 * Exposure closure of jubik/response.py:12-45 (``exposures * x`` with
 * numpy broadcasting of operands whose shapes are suffixes of the output:
 * out[i] = e[i % n_e] * x[i % n_x], i < n_out).  The exposure cut
 * (response.py:37-40) is applied on the host. */
int jubik_exposure_apply(int dtype, const void* exposure, int64_t n_e, const void* x,
                         int64_t n_x, void* out, int64_t n_out, void* stream);
/* Read-out of jubik/response.py:94-110 (``x[i][~mask[i]]``) as a gather by
 * the C-order flat indices of the kept pixels, and its transpose (scatter
 * into a zero array of n elements). */
int jubik_readout_gather(int dtype, const void* x, const int64_t* index, int64_t nnz, void* out,
                         void* stream);
int jubik_readout_scatter(int dtype, const void* v, const int64_t* index, int64_t nnz, int64_t n,
                          void* out, void* stream);
/* Mock counts: counts[i] ~ Poisson(lam[i]) for n device values, as the
 * mock-data generator draws them per read-out key
 * (jubik/data.py:139-145, jax.random.poisson).  A draw is a pure function of
 * (seed, i, lam[i]); the generator differs from JAX's, so the counts match
 * the reference in distribution, not bit for bit. */
int jubik_poisson_sample(int dtype, const void* lam, int64_t n, uint64_t seed, int32_t* counts,
                         void* stream);
/* Poisson NLL of nifty.re's Poissonian(data) as composed at
 * jubik/instruments/chandra/chandra_likelihood.py:37 and
 * jubik/instruments/erosita/erosita_likelihood.py:89:
 * nll (device double) = sum(lam - d log lam); grad = 1 - d / lam (may be
 * NULL).  ``work``: JUBIK_POISSON_WORK_BYTES of device memory. */
int jubik_poisson_nll_grad(int dtype, const void* lam, const void* data, int64_t n, double* nll,
                           void* grad, void* work, void* stream);
/* jubik_poisson_nll_grad without a workspace argument and without any
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The FP64 partial sums live in a
 * static device array of JUBIK_POISSON_WORK_BYTES (1024 partials, one per
 * reduction block; any n is reduced in at most 1024 blocks) that each GPU
 * holds until the process ends.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (concurrent streams or host threads) race
 * and must be serialised by the caller. */
int jubik_poisson_nll_grad_noalloc(int dtype, const void* lam, const void* data, int64_t n,
                                   double* nll, void* grad, void* stream);
/* Fisher metric of that likelihood in data space, applied to t:
 * out = t / lam (sqrt_form = 0, nifty.re Poissonian.metric) or its left
 * square root out = t / sqrt(lam) (sqrt_form = 1, left_sqrt_metric), the
 * two products nifty.re's sampling calls on the likelihood composed at
 * jubik/instruments/chandra/chandra_likelihood.py:37. */
int jubik_poisson_metric(int dtype, const void* lam, const void* t, int64_t n, int sqrt_form,
                         void* out, void* stream);

/* ---------------------------------------------------------------- DFT comparison
 * Batched 128-point complex DFTs (FP32 complex in and out, B vectors, B a
 * multiple of 256), no counterpart in the reference: method 0 the FP32 FFT
 * the PSF kernels use, method 1 a split-BF16 DFT as real GEMMs on the
 * matrix cores (each product as hi*hi + hi*lo + lo*hi, FP32 accumulation),
 * the comparison asked for the 2048^2 x 4 energy configuration. */
typedef struct jubik_dft128 jubik_dft128;
int jubik_dft128_create(jubik_dft128** plan);
int jubik_dft128_run(jubik_dft128* plan, int method, const void* x, void* y, int64_t B,
                     void* stream);
int jubik_dft128_destroy(jubik_dft128* plan);

/* ---------------------------------------------------------------- PSF preparation
 * eROSITA patch PSFs of
 * jubik/instruments/erosita/erosita_psf_utils.py:185-264 (psf_interpolator,
 * with get_psf :90-164) for one energy: the nrad tabulated PSFs (device,
 * nrad x ph x pw) at off-axis radii rs are weighted linearly by each patch
 * centre's radius, rotated by its azimuth about pointing_center and sampled
 * bilinearly (map_coordinates order 1, mode 'nearest').  rs,
 * patch_center_ids (nrad x 2), patch_deltas (2) and pointing_center (2) are
 * host arrays.  Writes slot e of out (P, E, ...): layout 0 the (H, W) frame
 * rolled to the origin (the reference's output), layout 1 the centred
 * 2H/npatch-wide window linpatch_convolve cuts from it.  Set-up routine: it
 * allocates a small table and synchronises the stream. */
int jubik_psf_interpolate(int dtype, const void* psfs, int64_t nrad, int64_t ph, int64_t pw,
                          const double* rs, const double* patch_center_ids,
                          const double* patch_deltas, const double* pointing_center, int64_t H,
                          int64_t W, double dH, double dW, int64_t npatch, int64_t e, int64_t E,
                          int layout, void* out, void* stream);

/* ---------------------------------------------------------------- GP sky
 * Correlated-field sky of
 * jubik/sky_model/multifrequency/spectral_product_mf_sky.py:171-298 (and
 * the diffuse exp(CF) of jubik/sky_models.py:306-341) with a fixed
 * amplitude field amp = hdvol * fluct * A[pd] (A[0] = 0), on an N x N GP
 * grid cropped at the origin to the (H, W) sky:
 *   s_e = exp(zm + HT(amp_ref xi_ref) + (HT(amp_alpha xi_alpha) + mu) dnu_e)
 * HT = Re(FFT) + Im(FFT) (nifty.re hartley).  has_alpha = 0 gives the
 * single-frequency sky exp(zm + HT(amp_ref xi_ref)).                      */
typedef struct jubik_gp jubik_gp;

int jubik_gp_create(jubik_gp** plan, int dtype, int64_t N, int64_t H, int64_t W, int64_t E,
                    int has_alpha);
int jubik_gp_destroy(jubik_gp* plan);
/* sky (E, H, W).  dnu: device array of E relative log-frequencies. */
int jubik_gp_forward(jubik_gp* plan, const void* amp_ref, const void* amp_alpha,
                     const void* xi_ref, const void* xi_alpha, double zm, double mu,
                     const void* dnu, void* sky, void* stream);
/* VJP: given sky (E, H, W) and its cotangent, the cotangents of xi_ref and
 * xi_alpha (N x N each; g_alpha ignored when has_alpha = 0). */
int jubik_gp_vjp(jubik_gp* plan, const void* amp_ref, const void* amp_alpha, const void* dnu,
                 const void* sky, const void* cot_sky, void* g_ref, void* g_alpha, void* stream);

/* ---------------------------------------------------------------- fused model
 * lam_M = Exposure_M(linpatch(a * s(xi), K_M)), read out where counts >= 0,
 * NLL = sum (lam - d log lam): the likelihood of
 * jubik/instruments/erosita/erosita_likelihood.py:73-89 and
 * jubik/instruments/chandra/chandra_response.py:136-142 composed with the
 * sky.  ``exposure`` (M, E, H, W) includes the effective area and is zero
 * where cut; ``counts`` (M, E, H, W) int32 holds the data, -1 where masked.
 * The model keeps pointers to the gp and linpatch plans and to the caller's
 * exposure/counts buffers.                                                  */
typedef struct jubik_model jubik_model;

int jubik_model_create(jubik_model** model, jubik_gp* gp, jubik_linpatch* lp, double pixel_area);
/* A model that evaluates up to max_samples positions per call: the VI
 * samples of nifty.re's optimize_kl (demos/x-ray-imaging.py:404-430, sample
 * counts from jubik/minimization_parser.py:240-287), at each of which the
 * likelihood gradient and metric are evaluated.  It owns workspaces for
 * max_samples sets of jobs. */
int jubik_model_create_batched(jubik_model** model, jubik_gp* gp, jubik_linpatch* lp,
                               double pixel_area, int64_t max_samples);
int jubik_model_destroy(jubik_model* model);
int jubik_model_set_data(jubik_model* model, const void* exposure, const int32_t* counts);
/* One gradient evaluation: nll (device double) and the gradient with
 * respect to (xi_ref, xi_alpha).  Forward, Poisson residual and the
 * hand-written adjoint chain; the sky and lam of this position stay in the
 * model (for jubik_model_metric). */
int jubik_model_grad(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                     const void* xi_ref, const void* xi_alpha, double zm, double mu,
                     const void* dnu, double* nll, void* g_ref, void* g_alpha, void* stream);
/* jubik_model_grad plus the gradients with respect to the sky model's
 * amplitude fields and offsets, the inputs of the reference's amplitude
 * models and zero-mode priors (sky composition
 * jubik/sky_model/multifrequency/spectral_product_mf_sky.py:171-298; the
 * amplitude models of spectral_product_utils/normalized_amplitude_model.py:61-139
 * produce amp_ref / amp_alpha): g_amp_ref = xi_ref * HT(u_ref) and
 * g_amp_alpha = xi_alpha * HT(u_alpha) (N x N, model dtype), g_offsets =
 * (d NLL / d zm, d NLL / d mu_alpha) (two values, model dtype), where u is
 * the cotangent of log s.  Sets the metric base point like
 * jubik_model_grad. */
int jubik_model_grad_hyper(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                           const void* xi_ref, const void* xi_alpha, double zm, double mu,
                           const void* dnu, double* nll, void* g_ref, void* g_alpha,
                           void* g_amp_ref, void* g_amp_alpha, void* g_offsets, void* stream);
/* MGVI metric product J^T diag(1/lam) J v at the position of the last
 * jubik_model_grad call (the metric nifty.re's sampling uses).  The base
 * point holds the sky and lam of that call: pass the same amplitude fields
 * and dnu, and after the amplitude fields or offsets change evaluate the
 * gradient again before the next metric product. */
int jubik_model_metric(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                       const void* dnu, const void* v_ref, const void* v_alpha, void* out_ref,
                       void* out_alpha, void* stream);
/* Domain check of the last gradient evaluation (single or batch): the
 * evaluation counts read-out pixels (counts >= 0) where lam <= 0 or is NaN,
 * where log lam is undefined and the NLL becomes inf or NaN, as the
 * reference's does (demos/x-ray-imaging.py:312-315: zero-exposure pixels
 * must be masked).  Synchronises ``stream``; writes the count per sample to
 * counts[0 .. S) when counts is not NULL and returns JUBIK_EDOMAIN with a
 * message when any is positive. */
int jubik_model_check_domain(jubik_model* model, int64_t* counts, void* stream);
/* S gradient evaluations in one pass: xi_ref, xi_alpha, g_ref, g_alpha are
 * (S, N, N) stacks, nll holds S device doubles.  The metric base point is
 * not changed. */
int jubik_model_grad_batch(jubik_model* model, int64_t S, const void* amp_ref,
                           const void* amp_alpha, const void* xi_ref, const void* xi_alpha,
                           double zm, double mu, const void* dnu, double* nll, void* g_ref,
                           void* g_alpha, void* stream);
/* Point sources added to the diffuse sky: jubik/sky_models.py:376-441 (the
 * SkyModel point-source component, sky = diffuse + points at 202) and
 * jubik/sky_model/multifrequency/spectral_product_mf_point_sources.py:99-287
 * (MultiFrequencyInvGammaSky without spectral deviations).  Per pixel
 *   p_e = I(xi_ps) exp(alpha dnu_e),  alpha = alpha_mean + alpha_std xi_ps_alpha,
 * I = invgamma_prior(a, q)(xi_ps): the inverse-gamma variable of shape a and
 * scale q with the quantile of the standard normal xi_ps.
 * jubik_model_set_point_sources sets the priors (enable = 0 switches them
 * off) and allocates the model's point-source buffers.
 * jubik_model_grad_points evaluates S positions (xi_ps, xi_ps_alpha, g_ps,
 * g_ps_alpha are (S, H, W) stacks on the sky grid, the rest as in
 * jubik_model_grad_batch): S = 1 sets the metric base point like
 * jubik_model_grad, S > 1 uses the batch slots like jubik_model_grad_batch.
 * The metric entry points then apply to the diffuse latents at lam of the
 * total sky.  Without a spectral index (single-energy sky) xi_ps_alpha and
 * g_ps_alpha may be NULL and p = I. */
int jubik_model_set_point_sources(jubik_model* model, int enable, double a, double q,
                                  double alpha_mean, double alpha_std);
int jubik_model_grad_points(jubik_model* model, int64_t S, const void* amp_ref,
                            const void* amp_alpha, const void* xi_ref, const void* xi_alpha,
                            double zm, double mu, const void* dnu, const void* xi_ps,
                            const void* xi_ps_alpha, double* nll, void* g_ref, void* g_alpha,
                            void* g_ps, void* g_ps_alpha, void* stream);
/* The inverse-gamma prior transform of nifty.re's invgamma_prior(a, q) as
 * used at jubik/sky_models.py:386-391: log_i[k] = log I(xi[k]) and (if not
 * NULL) dlog_i[k] = d log I / d xi at n standard-normal values xi (device
 * arrays of the dtype). */
int jubik_invgamma_prior(int dtype, const void* xi, int64_t n, double a, double q, void* log_i,
                         void* dlog_i, void* stream);
/* geoVI nonlinear sample updates (nifty.re's nonlinearly_update, the
 * "nonlinear_resample" / "nonlinear_update" sample modes of
 * demos/erosita_demo.py:304-316 and demos/configs/eROSITA_demo.yaml:148-166):
 * a sample x near the base point p (the last jubik_model_grad) is moved to
 * minimise 1/2 |r(x)|^2 with
 *   r(x) = (x - p) + J(p)^T D_p 2 (sqrt lam(x) - sqrt lam(p)) - m,
 * D = diag(lam^-1/2) on read-out pixels and m the metric sample of the
 * linear draw (2 sqrt lam is the Poisson likelihood's transformation, in
 * whose coordinates its Fisher metric is the identity).  Its Gauss-Newton
 * metric is (I + K)^T (I + K), K = J(p)^T D_p D_x J(x).
 * jubik_model_geovi_point evaluates the model at x (into the batch slot 1,
 * the base point stays) and prepares the scalings; nll receives the NLL at
 * x (one device double).  jubik_model_geovi_apply: mode 0 out = K v, mode
 * 1 out = K^T v, mode 2 out = J(p)^T D_p 2 (sqrt lam(x) - sqrt lam(p)) (v
 * unused), mode 3 out = (K + K^T + K^T K) v.  jubik_model_geovi_solve: CG on
 * (I + K + K^T + K^T K) x = b, arguments as jubik_model_metric_solve.  A
 * gradient evaluation moves the points: call jubik_model_geovi_point again
 * (JUBIK_ENOTREADY otherwise).  No counterpart in the reference beyond the
 * nifty.re call it configures. */
int jubik_model_geovi_point(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                            const void* x_ref, const void* x_alpha, double zm, double mu,
                            const void* dnu, double* nll, void* stream);
int jubik_model_geovi_apply(jubik_model* model, int mode, const void* amp_ref,
                            const void* amp_alpha, const void* dnu, const void* v_ref,
                            const void* v_alpha, void* out_ref, void* out_alpha, void* stream);
int jubik_model_geovi_solve(jubik_model* model, const void* amp_ref, const void* amp_alpha,
                            const void* dnu, const void* b_ref, const void* b_alpha, void* x_ref,
                            void* x_alpha, int maxiter, double rtol, double absdelta,
                            int check_every, double* info, void* stream);
/* S metric products at the base point in one pass: v and out are (S, N, N)
 * stacks (the S conjugate-gradient systems of an MGVI sample draw). */
int jubik_model_metric_batch(jubik_model* model, int64_t S, const void* amp_ref,
                             const void* amp_alpha, const void* dnu, const void* v_ref,
                             const void* v_alpha, void* out_ref, void* out_alpha, void* stream);
/* The tiles of the model's own evaluations (S samples per pass), as for
 * jubik_linpatch_autotune; a divisor of 0 in jubik_model_set_tile follows
 * the plan's tile again. */
int jubik_model_set_tile(jubik_model* model, int col_div, int row_div);
int jubik_model_autotune(jubik_model* model, int64_t S, int reps, int* col_div, int* row_div,
                         double* ms, void* stream);
/* Graph capture and launch (no counterpart in the reference, which relies
 * on XLA's compilation of the jitted likelihood): the gradient evaluation of
 * jubik_model_grad (S = 0) or jubik_model_grad_batch (S >= 1) with these
 * fixed buffers, captured once into an executable HIP graph.  Each launch
 * reads the buffers' current contents; launches of an S = 0 graph also set
 * the metric base point.  Capture fails while kernel timing is enabled. */
/* Lifetime: a graph refers to the models, plans and buffers it captured;
 * all of them must outlive the graph (destroy the graph first).  zm and mu
 * are fixed at capture; the amplitude fields and positions are read from
 * their buffers at every launch. */
typedef struct jubik_graph jubik_graph;
int jubik_model_grad_graph(jubik_graph** graph, jubik_model* model, int64_t S,
                           const void* amp_ref, const void* amp_alpha, const void* xi_ref,
                           const void* xi_alpha, double zm, double mu, const void* dnu,
                           double* nll, void* g_ref, void* g_alpha);
int jubik_graph_launch(jubik_graph* graph, void* stream);
int jubik_graph_destroy(jubik_graph* graph);
/* Several gradient evaluations in one graph: the modules, energies and
 * samples of observations on uneven grids (the instruments' per-module
 * vmaps, chandra_response.py:127-130 and erosita_response.py:209-212, with
 * one model per grid).  Each descriptor is one jubik_model_grad (samples =
 * 0, sets that model's base point at launch) or jubik_model_grad_batch
 * (samples = S) call; the evaluations become parallel branches of the
 * graph, so small grids that cannot fill the device alone run side by side.
 * Descriptors may not share a model or a plan; at most 64. */
typedef struct jubik_grad_desc {
  jubik_model* model;
  int64_t samples;
  const void* amp_ref;
  const void* amp_alpha;
  const void* xi_ref;
  const void* xi_alpha;
  double zm;
  double mu;
  const void* dnu;
  double* nll;
  void* g_ref;
  void* g_alpha;
} jubik_grad_desc;
int jubik_graph_capture_batch(jubik_graph** graph, const jubik_grad_desc* descs, int64_t count);
/* Metric products of several models in one graph: each descriptor is one
 * jubik_model_metric_batch call (S tangents at that model's base point),
 * captured as parallel branches, e.g. the S = 32 antithetic metric products
 * of three observations on uneven grids per CG iteration of a sample draw
 * (demos/x-ray-imaging.py:416-419).  Descriptors may not share a model or
 * a plan; at most 64; each model needs a base point at capture and keeps
 * the one it has at launch. */
typedef struct jubik_metric_desc {
  jubik_model* model;
  int64_t samples;
  const void* amp_ref;
  const void* amp_alpha;
  const void* dnu;
  const void* v_ref;
  const void* v_alpha;
  void* out_ref;
  void* out_alpha;
} jubik_metric_desc;
int jubik_graph_capture_metric_batch(jubik_graph** graph, const jubik_metric_desc* descs,
                                     int64_t count);
/* Left square root of the likelihood metric at the base point, for S noise
 * vectors: out = J^T diag(lam^-1/2) noise, noise (S, M, E, H, W) in data
 * space (ignored where counts < 0).  With noise ~ N(0, 1) this draws from
 * N(0, J^T diag(1/lam) J): the likelihood part of nifty.re's linear sample
 * draw (draw_linear_residual, called through optimize_kl at
 * demos/x-ray-imaging.py:404-430). */
int jubik_model_sqrt_metric_batch(jubik_model* model, int64_t S, const void* amp_ref,
                                  const void* amp_alpha, const void* dnu, const void* noise,
                                  void* out_ref, void* out_alpha, void* stream);
/* The KL metric's likelihood part at the S sample positions of the last
 * jubik_model_grad_batch: out = sum_s J_s^T diag(1/lam_s) J_s v for one
 * tangent v (N x N fields), each J_s and lam_s at its own sample position
 * (the metric nifty.re's optimize_kl minimises the KL with, Newton-CG
 * minimizer "M", demos/x-ray-imaging.py:421-425).  Divide by S and add v
 * for the KL metric of the standardised latents. */
int jubik_model_metric_samples(jubik_model* model, int64_t S, const void* amp_ref,
                               const void* amp_alpha, const void* dnu, const void* v_ref,
                               const void* v_alpha, void* out_ref, void* out_alpha,
                               void* stream);
/* Device-resident conjugate gradients (the CG of nifty.re's sample draws
 * and Newton steps, demos/x-ray-imaging.py:416-425): every scalar stays in
 * device memory, dot products are FP64 reductions in a fixed order, and an
 * iteration is a fixed sequence of kernel launches with no host
 * synchronisation, so a solve can be captured in a graph when
 * check_every = 0 (then exactly maxiter iterations are launched; systems
 * that have met their stopping rule take steps of length 0).  With
 * check_every = k > 0 the stream is synchronised every k iterations to stop
 * early once every system has converged.  A system stops when
 * |r| <= rtol |b|, or (absdelta > 0) when the energy 1/2 x.A.x - b.x
 * changes by at most absdelta in one iteration (cg_kwargs absdelta of the
 * demo).  info (device, may be NULL) receives 2 doubles per system:
 * iterations and |r| / |b|.  x may not alias b.  The solver workspace (four
 * vectors of max_samples x 2 x N^2 elements and the scalars) is allocated
 * once by jubik_model_reserve_solver, a set-up call.
 * jubik_model_metric_solve: S systems (I + J^T diag(1/lam) J) x_s = b_s at
 * the base point (the linear sample draws).  jubik_model_kl_solve: one
 * system (I + (1/S) sum_s J_s^T diag(1/lam_s) J_s) x = b at the S positions
 * of the last jubik_model_grad_batch (the Newton step of the KL). */
int jubik_model_reserve_solver(jubik_model* model);
int jubik_model_metric_solve(jubik_model* model, int64_t S, const void* amp_ref,
                             const void* amp_alpha, const void* dnu, const void* b_ref,
                             const void* b_alpha, void* x_ref, void* x_alpha, int maxiter,
                             double rtol, double absdelta, int check_every, double* info,
                             void* stream);
int jubik_model_kl_solve(jubik_model* model, int64_t S, const void* amp_ref,
                         const void* amp_alpha, const void* dnu, const void* b_ref,
                         const void* b_alpha, void* x_ref, void* x_alpha, int maxiter,
                         double rtol, double absdelta, int check_every, double* info,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif /* JUBIK_HIP_H */
