/*
 * tdtm -- MI355X-native 2D template matching (2DTM) search engine, C ABI.
 *
 * Plain pointers and sizes only.  All device pointers are HIP device allocations
 * (e.g. torch tensors' data_ptr()) on the current device; `stream` is a hipStream_t
 * (NULL = default stream).  Every entry point returns 0 on success or a non-zero
 * code; tdtm_last_error() then returns a thread-local message.  Work is enqueued
 * asynchronously on `stream`.
 *
 * Layouts (row-major, C order):
 *   image_dft     complex64 (H, W/2+1)       rfft of the pre-processed micrograph, DC = 0
 *   template_dft  complex64 (h, h, h/2+1)    3D rfft of the template, fftshifted in dims 0,1
 *   ctf_filters   float32   (C, D, h, h/2+1)
 *   whitening     float32   (h, h/2+1)
 *   euler_deg     float32   (O, 3)           (phi, theta, psi), ZYZ intrinsic, degrees
 *   statistics    (H, W) each: mip f32, best_index i32, sum_hi/sum_lo f32 (compensated
 *                 pair), sq_hi/sq_lo f32 (compensated pair)
 */
#ifndef TDTM_H
#define TDTM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdtm_match_plan* tdtm_match_plan_t;

/* Phases of a search step reported by tdtm_match_plan_phase_times. */
#define TDTM_NUM_PHASES 4 /* project (slice + filter + irfft2 + normalise) | rowfwd | colpass | rowpass */

/* Version of the ABI (bumped on any signature change). */
int tdtm_abi_version(void);

/* Thread-local message for the last failing call on this thread. */
const char* tdtm_last_error(void);

/*
 * Plan for a whole-orientation search: buffers and twiddle tables for an H x W image,
 * an h^3 template and up to max_orientation_batch orientations x num_defocus x
 * num_pixel correlations per step.  Replaces the per-device set-up inside
 * _core_match_template_single_gpu (src/leopard_em/backend/core_match_template.py:369-486).
 * Supported sizes: H in {64 .. 4096}, W in {64 .. 8192}, powers of two; h in {32 .. 512}
 * powers of two, h <= min(H, W).  Other image sizes run in the next power-of-two grid
 * (tdtm_match_plan_set_image_size, tdtm_irfft2_any).
 */
int tdtm_match_plan_create(int image_h, int image_w, int template_h, int max_orientation_batch,
                           int num_defocus, int num_pixel, tdtm_match_plan_t* out_plan);
int tdtm_match_plan_destroy(tdtm_match_plan_t plan);
/* Device bytes held by the plan. */
int64_t tdtm_match_plan_bytes(tdtm_match_plan_t plan);

/*
 * Image inside a larger FFT grid: the plan's H x W grid holds an image_h x image_w
 * micrograph zero-padded at the bottom / right (sizes that are not powers of two, e.g.
 * 4092 x 5760 in a 4096 x 8192 grid).  Sets the image size the projection normalisation
 * uses (normalize_template_projection's relative size and padded variance,
 * src/leopard_em/backend/utils.py:106-188); correlations at positions whose template
 * footprint lies inside the image are then those of the unpadded search.  Default: the
 * grid size.  A change releases the plan's captured step and refinement graphs (they are
 * re-captured on the next run).  No counterpart in the reference (torch.fft handles any
 * size there).
 */
int tdtm_match_plan_set_image_size(tdtm_match_plan_t plan, int image_h, int image_w);

/*
 * Anisotropic magnification (host doubles {a00, a01, a10, a11}, NULL = none): every central
 * slice is resampled bilinearly at A k before the filters (torch_fourier_slice's
 * transform_slice_2d at src/leopard_em/backend/cross_correlation.py:81-88, :231-238), and the
 * per-particle CTF of refinement is evaluated at A k (the mag_matrix entry of ctf_kwargs,
 * utils/ctf_utils.py:66-100).  A acts on (x, y) frequency column vectors; slice points whose
 * transformed x is negative take the conjugate of the sample at -A k.  Default: none.
 */
int tdtm_match_plan_set_magnification(tdtm_match_plan_t plan, const double* mag_matrix);

/*
 * Launch-shape options of a plan (no counterpart in the reference; the autotuner in
 * leopard_em_amd.backend.autotune sets them).  Keys: "colpass_columns" (1 or 2 columns
 * per column-pass block), "colpass_pair" (0/1: two correlations per column-pass
 * iteration sharing twiddles and barriers; default 0), "colpass_chunks" (correlation chunks
 * per column group, 0 = automatic), "colpass_twiddles" (0: global table, 1: LDS copy
 * for the four-step column transforms where it fits, 2: each lane's twiddles built once
 * in registers for all its correlations, default; one correlation per iteration,
 * otherwise the LDS copy), "colpass_store" (0: direct 8-byte stores; 1: rows
 * through LDS, 16-byte stores of two rows of one column of a tile, default), "colpass_persistent" (0: one
 * block per tile column and correlation chunk; 1: as many blocks as are resident at once,
 * each group of blocks writing the same lines taking an equal share of the (tile-column
 * pair, correlation) items, default; one correlation per iteration only), "rowpass_staging" (0: synchronous de-interleaved,
 * 1: asynchronous raw double buffer, 2: asynchronous raw copy + de-interleaving reformat,
 * 3: synchronous de-interleaved with batched loads; default 3), "rowpass_twiddles" (0: global
 * table, 1: LDS copy, default; needs staging 0 or 3), "rowpass_rows" (rows of a 4-row tile
 * row per row-pass block: 1 (default), 2 or 4; below 4 needs staging 0 or 3; staging 1 / 2
 * need 256 <= W <= 4096), "graph" (0/1: tdtm_match_run replays full-batch steps from a
 * captured HIP graph; default 1), "phase_timing" (0/1: record HIP events between the
 * phases of every step; tdtm_match_run then launches steps directly).
 * Results do not depend on the options.
 */
int tdtm_match_plan_set_option(tdtm_match_plan_t plan, const char* key, int value);

/*
 * With "phase_timing" on: waits for the recorded steps, adds up each phase's device
 * time over them into out_ms[TDTM_NUM_PHASES] (milliseconds), stores the number of
 * steps in *out_steps and releases the events.  No counterpart in the reference (its
 * benchmark times whole calls, benchmark/benchmark_match_template.py:88-123); used by
 * the benchmark and the launch-shape autotuner.
 */
int tdtm_match_plan_phase_times(tdtm_match_plan_t plan, double* out_ms, int* out_steps);

/*
 * Load the pre-processed image spectrum (re-tiled and scaled by 1/(H W) inside the
 * plan).  Replaces image_dft.to(device) at core_match_template.py:448.
 */
int tdtm_match_set_image(tdtm_match_plan_t plan, const void* image_dft, void* stream);

/*
 * projective_filters = ctf_filters * whitening (core_match_template.py:274), kept in
 * the plan.
 */
int tdtm_match_set_filters(tdtm_match_plan_t plan, const float* ctf_filters, const float* whitening,
                           void* stream);

/*
 * mip = -inf, best_index = -1, sums = 0 over n pixels
 * (core_match_template.py:472-486).
 */
int tdtm_stats_init(int64_t n, float* mip, int32_t* best_index, float* sum_hi, float* sum_lo,
                    float* sq_hi, float* sq_lo, void* stream);

/*
 * One orientation batch [start, start+count) of the search: Fourier slice extraction,
 * projective filter, small irfft2, projection normalisation, padded FFT
 * cross-correlation against the image and the MIP / argmax / sum / sum-of-squares
 * update, fused so correlation maps never reach HBM.  Replaces one iteration of the
 * loop at core_match_template.py:504-550, i.e.
 * do_streamed_orientation_cross_correlate (cross_correlation.py:13-154) followed by
 * do_iteration_statistics_updates (utils.py:191-256).  Global search index of
 * correlation (c, d, o) is c*D*O + d*O + o with O = num_orientations.
 */
int tdtm_match_step(tdtm_match_plan_t plan, const void* template_dft, const float* euler_deg,
                    int num_orientations, int start, int count, float* mip, int32_t* best_index,
                    float* sum_hi, float* sum_lo, float* sq_hi, float* sq_lo, void* stream);

/*
 * Orientations [start, end) in steps of max_orientation_batch: the orientation loop of
 * _core_match_template_single_gpu (core_match_template.py:492-554) on one device, each
 * step as tdtm_match_step.  With the plan option "graph" (default 1) the full-batch steps
 * are replayed from a HIP graph: one step is captured once (its first orientation read
 * from device memory and advanced on the device) and launched once per batch; the graph
 * is re-captured when the template, Euler or statistics pointers or num_orientations
 * change.  A last partial batch runs as a plain step.  Results are bit-identical to
 * calling tdtm_match_step per batch.
 */
int tdtm_match_run(tdtm_match_plan_t plan, const void* template_dft, const float* euler_deg,
                   int num_orientations, int start, int end, float* mip, int32_t* best_index,
                   float* sum_hi, float* sum_lo, float* sq_hi, float* sq_lo, void* stream);

/*
 * Final statistics (FP64): mean = sum/N, std = sqrt(max(sumsq/N - mean^2, 0)),
 * scaled_mip = (mip - mean)/std (0 where std == 0), and best index decoded to
 * (phi, theta, psi, defocus).  Replaces decode_global_search_index and scale_mip
 * (process_results.py:67-183).  Output maps are float32 (n pixels).
 */
int tdtm_match_finalize(int64_t n, int64_t total_positions, const float* mip, const int32_t* best_index,
                        const float* sum_hi, const float* sum_lo, const float* sq_hi, const float* sq_lo,
                        const float* euler_deg, int num_orientations, const float* defocus_values,
                        int num_defocus, float* out_scaled_mip, float* out_mean, float* out_std,
                        float* out_phi, float* out_theta, float* out_psi, float* out_defocus,
                        void* stream);

/*
 * Cross-device merge of partial statistics (packed max with first-operand precedence
 * on ties, compensated sums): a <- merge(a, b).  Replaces
 * aggregate_distributed_results (process_results.py:7-64) for device-resident maps.
 */
int tdtm_stats_merge(int64_t n, float* mip_a, int32_t* idx_a, float* sum_hi_a, float* sum_lo_a,
                     float* sq_hi_a, float* sq_lo_a, const float* mip_b, const int32_t* idx_b,
                     const float* sum_hi_b, const float* sum_lo_b, const float* sq_hi_b,
                     const float* sq_lo_b, void* stream);

/*
 * Local refinement of one particle: for every orientation offset (composed with the
 * particle's pose as R(offset) @ R(pose)), defocus offset and pixel-size offset,
 * correlate the template against the particle crop, z-score the valid region with
 * the supplied mean / std maps and keep the best (first strict maximum over batches of
 * `max_orientation_batch` offsets).  Replaces _core_refine_template_single_thread
 * (src/leopard_em/backend/core_refine_template.py:555-767) including its per-particle
 * CTF stack (utils/ctf_utils.py:19-116) and the angle composition / wrap of
 * core_refine_template.py:504-536.  The plan is created with the crop size as image
 * size (or, for crop sizes that are not powers of two, with the next power-of-two grid
 * and tdtm_match_plan_set_image_size(crop size); particle_dft is then the spectrum of the
 * zero-padded crop, e.g. from tdtm_irfft2_any + tdtm_rfft2), num_defocus =
 * len(defocus_offsets), num_pixel = len(pixel_offsets).
 *   particle_dft      complex64 (grid H, grid W/2+1), euler_pose float32 (3) on the device,
 *   offsets           float32 (n_off, 3) degrees,  defocus_offsets (D), pixel_offsets (C),
 *   ctf_params        host doubles: voltage kV, Cs mm, amplitude contrast, phase shift
 *                     deg, pixel size A, B-factor A^2,
 *   projective_filter float32 (h, h/2+1),  corr_mean / corr_std float32 (crop H-h+1, crop W-h+1),
 *   out               float32[14] on the device: max_cc, max_z, phi, theta, psi (composed,
 *                     wrapped), defocus offset, pixel-size offset, pos_y, pos_x (valid-crop
 *                     coordinates), angle_idx, unused, phi/theta/psi offsets.
 */
int tdtm_refine_particle(tdtm_match_plan_t plan, const void* template_dft, const void* particle_dft,
                         const float* euler_pose, const float* offsets, int n_off,
                         const float* defocus_offsets, const float* pixel_offsets, double defocus_u,
                         double defocus_v, double defocus_angle, const double* ctf_params,
                         const float* projective_filter, const float* corr_mean, const float* corr_std,
                         float* out, void* stream);

/*
 * Refinement of a whole particle stack in one call: tdtm_refine_particle for every
 * particle i, each with its own grid -- orientation offsets offsets[off_start[i] ..
 * + off_count[i]) (float32 (., 3) on the device), defocus offsets defocus_offsets[d_start[i]
 * .. + d_count[i]) and pixel-size offsets pixel_offsets[c_start[i] .. + c_count[i])
 * (float32 on the device; d_count[i] <= the plan's num_defocus, c_count[i] <= its
 * num_pixel) -- its own defocus_u / v / angle (host double arrays) and projective filter,
 * mean and std maps.  Replaces the particle loop of core_refine_template
 * (src/leopard_em/backend/core_refine_template.py:206-327 splits the stack over devices,
 * :555-767 refines one particle) for shared grids and for uneven per-particle grids.
 *   particle_dft        complex64, particle i at particle_dft + i * particle_stride
 *                       (complex elements, >= grid H x (W/2+1)), resident on the device
 *   euler_poses         float32 (n, 3);  projective_filters float32 (n, h, h/2+1)
 *   corr_mean/corr_std  float32 (n, crop H-h+1, crop W-h+1);  out float32 (n, 14) as
 *                       tdtm_refine_particle
 *   off_start .. c_count host int arrays of n entries (starts index the device arrays)
 * With the plan option "graph" (default 1) the full batches of offsets are replayed from
 * HIP graphs captured once per (defocus count, pixel-size count): the first offset, the
 * result slot and the particle index are read from device memory and advanced by the
 * graph.  Results are identical to tdtm_refine_particle per particle.
 */
int tdtm_refine_batch(tdtm_match_plan_t plan, const void* template_dft, const void* particle_dft, int n_particles,
                      int64_t particle_stride, const float* euler_poses, const float* offsets, const int* off_start,
                      const int* off_count, const float* defocus_offsets, const int* d_start, const int* d_count,
                      const float* pixel_offsets, const int* c_start, const int* c_count, const double* defocus_u,
                      const double* defocus_v, const double* defocus_angle, const double* ctf_params,
                      const float* projective_filters, const float* corr_mean, const float* corr_std, float* out,
                      void* stream);

/* ------------------------------------------------------- orientation-proposal network
 * No counterpart in the reference.  A small CNN that scores coarse orientation bins (and
 * one "empty tile" class, the last) for 64 x 64 tiles (micrograph tiles of 64 s x 64 s pixels
 * averaged s x s), trained on the orientation maps of the plain search; leopard_em_amd.proposal
 * uses it to choose which orientations a tile's search visits.  Layers: four 3x3 conv +
 * ReLU + 2x2 average pooling (16, 32, 64, 64 channels), fc 1024 -> 256 + ReLU, fc 256 ->
 * num_classes.  Matrix products on bf16 MFMA with float32 accumulation; float32 master
 * weights.  Parameter vector layout: per conv layer W (Cout, 3, 3, Cin) then b (Cout); W_fc1
 * (256, 4, 4, 64) (inputs in (y, x, c) order) then b_fc1; W_fc2 (num_classes, 256) then
 * b_fc2.  Tiles float32 (batch, 64, 64); logits / targets float32 (batch, num_classes),
 * targets are probability distributions; all on the device.
 */
typedef struct tdtm_proposal_net* tdtm_proposal_net_t;
int tdtm_proposal_create(int num_classes, int max_batch, tdtm_proposal_net_t* out);
int tdtm_proposal_destroy(tdtm_proposal_net_t net);
int64_t tdtm_proposal_num_params(tdtm_proposal_net_t net);
/* Load parameters (float32, tdtm_proposal_num_params) and reset the optimiser state. */
int tdtm_proposal_set_params(tdtm_proposal_net_t net, const float* params, void* stream);
/* Copy out the parameters and / or the gradients of the last backward pass (NULL: skip). */
int tdtm_proposal_get_params(tdtm_proposal_net_t net, float* params, float* grads, void* stream);
int tdtm_proposal_forward(tdtm_proposal_net_t net, const float* tiles, int batch, float* logits, void* stream);
/* Softmax cross-entropy of the last forward's logits against targets: per-sample loss
 * (batch floats) and every parameter's gradient of the batch-mean loss. */
int tdtm_proposal_backward(tdtm_proposal_net_t net, const float* targets, int batch, float* loss, void* stream);
int tdtm_proposal_adam_step(tdtm_proposal_net_t net, float lr, float beta1, float beta2, float eps, void* stream);

/* ---------------------------------------------------------------- input preparation
 * Once-per-search inputs built on the device.  The reference builds them on the host
 * in MatchTemplateManager.make_backend_core_function_kwargs
 * (src/leopard_em/pydantic_models/managers/match_template_manager.py:140-218).
 */

/*
 * Unnormalised forward 2D real FFT: image float32 (H, W) -> out complex64 (H, W/2+1),
 * as torch.fft.rfftn(image) (match_template_manager.py:160).  H, W powers of two,
 * H in [32, 4096], W in [32, 8192].
 */
int tdtm_rfft2(const float* image, int H, int W, void* out, void* stream);

/*
 * fftshift -> rfftn -> fftshift(dims 0, 1) of a cubic volume: volume float32 (n, n, n)
 * -> out complex64 (n, n, n/2+1).  Replaces volume_to_rfft_fourier_slice
 * (src/leopard_em/utils/image_processing.py:179-202).  n a power of two in [32, 512].
 */
int tdtm_volume_rfft(const float* volume, int n, void* out, void* stream);

/*
 * Inverse 2D real DFT of any size (direct sums in double): spectrum complex64 (H, W/2+1)
 * -> image float32 rows 0 .. H-1, columns 0 .. W-1 of a row-major array with row stride
 * ld (>= W) floats, with torch.fft.irfftn's conventions (1/(H W) normalisation, the
 * imaginary parts of DC and Nyquist ignored).  The other elements of `image` are not
 * written, so a zero-filled array receives the image zero-padded.  Used to move a
 * spectrum whose size is not a power of two onto the padded grid of a plan (once per
 * micrograph).  1 <= H <= 4096, W even <= 8192; workspace of
 * tdtm_irfft2_any_workspace_bytes(H, W) device bytes.  No counterpart in the reference.
 */
int64_t tdtm_irfft2_any_workspace_bytes(int H, int W);
int tdtm_irfft2_any(const void* spectrum, int H, int W, float* image, int64_t ld, void* workspace, void* stream);

/* Device workspace bytes tdtm_prepare_image needs (-1: unsupported size or bin count). */
int64_t tdtm_prepare_image_workspace_bytes(int H, int W, int num_freq_bins);

/*
 * Pre-processed image spectrum and template-grid whitening filter with the whitening
 * filter on or off and the bandpass / arbitrary-curve filters off:
 *   X = rfft2(image), X[0,0] = 0                      (match_template_manager.py:160-161)
 *   whitening: radial power of X in num_freq_bins bins on [0, max_freq] cycles / pixel
 *   (0 = min(H, W)/2 + 1 bins) -> curve 1/sqrt(P(f)), max-normalised, linearly
 *   interpolated onto an rfft grid, 1 beyond max_freq  (config/correlation_filters.py:49-108);
 *   all ones when whitening = 0
 *   image_dft = X * whitening(H, W) normalised to unit real-space variance
 *                                                     (utils/image_processing.py:16-75)
 *   whitening_template = whitening on the (h, h/2+1) grid  (match_template_manager.py:169-177)
 * image float32 (H, W), H <= 4096 by an even W <= 8192 (powers of two >= 32 take the
 * FFT kernels, other sizes direct double-precision DFTs); image_dft complex64
 * (H, W/2+1); whitening_template float32 (h, h/2+1); workspace of
 * tdtm_prepare_image_workspace_bytes(H, W, num_freq_bins) device bytes.
 */
int tdtm_prepare_image(const float* image, int H, int W, int h, int whitening, int num_freq_bins, double max_freq,
                       void* image_dft, float* whitening_template, void* workspace, void* stream);

/*
 * tdtm_prepare_image with the bandpass and arbitrary-curve filters of the pre-processing
 * (config/correlation_filters.py:164-334, combined at :374-410; phase randomisation is not
 * part of the combined filter there, :395-396).  Filters on |k| in cycles / pixel:
 *   bandpass (bandpass != 0): 1 on bp_low <= |k| <= bp_high, a raised-cosine edge
 *   0.5 (1 + cos(pi d / bp_falloff)) at distance 0 < d <= bp_falloff outside the band, 0
 *   beyond (hard edges for bp_falloff = 0);
 *   curve (n_curve > 0 points, host arrays, increasing frequencies, n_curve <= 64): linear
 *   interpolation of curve_amp at |k|, 1 below the first / above the last frequency.
 * The image spectrum is multiplied by whitening x bandpass x curve on its own grid, the
 * bandpass sets the dimensionality sum(bp) + sum(bp[:, 1:-1]) of the normalisation
 * (utils/image_processing.py:60-67, match_template_manager.py:163-165), and
 * template_filter = whitening x bandpass x curve on the (h, h/2+1) grid.  Workspace of
 * tdtm_prepare_image_filtered_workspace_bytes(H, W, num_freq_bins) device bytes.
 */
int64_t tdtm_prepare_image_filtered_workspace_bytes(int H, int W, int num_freq_bins);
int tdtm_prepare_image_filtered(const float* image, int H, int W, int h, int whitening, int num_freq_bins,
                                double max_freq, int bandpass, double bp_low, double bp_high, double bp_falloff,
                                int n_curve, const double* curve_freq, const double* curve_amp, void* image_dft,
                                float* template_filter, void* workspace, void* stream);

/*
 * tdtm_prepare_image without any device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same outputs as tdtm_prepare_image,
 * but the twiddle table and the workspace live in one static device array per GPU of at
 * most TDTM_PREPARE_NOALLOC_SCRATCH_BYTES, held until the process ends; the call makes no
 * hipMalloc / hipFree.  Limits: H and W powers of two in [32, 4096] x [32, 8192] (the
 * direct-DFT path of other sizes needs an image-sized workspace) and at most 8192
 * frequency bins; larger requests fail with an error.  Every call on a GPU uses that
 * GPU's one array: calls that overlap on one GPU (two streams, two threads) race.
 */
#define TDTM_PREPARE_NOALLOC_SCRATCH_BYTES (320 * 1024)
int tdtm_prepare_image_noalloc(const float* image, int H, int W, int h, int whitening, int num_freq_bins,
                               double max_freq, void* image_dft, float* whitening_template, void* stream);

/*
 * CTF x B-envelope filter stack, out float32 (num_pixel, num_defocus, h, h/2+1):
 * replaces calculate_ctf_filter_stack_full_args (src/leopard_em/utils/ctf_utils.py:19-116)
 * with torch_ctf's sign convention (CTF(0) = +amplitude contrast).  defocus = offsets +
 * (u + v)/2, astigmatism |u - v|/2 at defocus_angle degrees, Cs scaled per pixel size
 * as Cs / ((p + dp)/p)^4 (utils/search_utils.py:27-28).  defocus_offsets (num_defocus)
 * and pixel_offsets (num_pixel) float32 on the device; ctf_params host doubles as for
 * tdtm_refine_particle.
 */
int tdtm_ctf_filter_stack(int h, int num_defocus, int num_pixel, double defocus_u, double defocus_v,
                          double defocus_angle, const float* defocus_offsets, const float* pixel_offsets,
                          const double* ctf_params, float* out, void* stream);

/*
 * tdtm_ctf_filter_stack with an anisotropic magnification matrix (host doubles
 * {a00, a01, a10, a11}, NULL = none): the CTF at rfft frequency k = (kx, ky) is evaluated at
 * A k, A acting on (x, y) column vectors; the B-envelope stays at k.  Replaces the
 * transform_matrix argument of calculate_ctf_2d (src/leopard_em/utils/ctf_utils.py:66-100,
 * OpticsGroup.mag_matrix_tensor).
 */
int tdtm_ctf_filter_stack_mag(int h, int num_defocus, int num_pixel, double defocus_u, double defocus_v,
                              double defocus_angle, const float* defocus_offsets, const float* pixel_offsets,
                              const double* ctf_params, const double* mag_matrix, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TDTM_H */
