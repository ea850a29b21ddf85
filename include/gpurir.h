/*
 * gpurir.h -- C ABI of libgpurir, the MI355X (gfx950) room-impulse-response engine.
 *
 * Plain C types only.  Every function returns 0 on success or a negative code
 * (-1 invalid argument, -2 HIP runtime error, -3 out of memory) and then leaves
 * a message readable with gpurir_last_error() (thread-local).  Nothing here
 * exits the process.
 *
 * Pointer conventions: "_device" entry points take device pointers and a
 * hipStream_t passed as void* (NULL = default stream) and return as soon as the
 * work is enqueued; the plain entry points take host pointers and return when
 * the result is in host memory.  All outputs are caller-allocated.
 *
 * Calls share one grow-only device workspace per device: calls that overlap on
 * different streams of one device must be ordered by the caller.
 *
 * Reference: DavidDiazGuerra/gpuRIR v1.2.0 (src/gpuRIR_cuda.h, src/gpuRIR_cuda.cu,
 * src/python_bind.cpp).  Each declaration names the reference interface it replaces.
 */
#ifndef GPURIR_H
#define GPURIR_H

#ifdef __cplusplus
extern "C" {
#endif

/* Polar patterns: src/gpuRIR_cuda.h:3-9 (DIR_OMNI ... DIR_BIDIR), gpuRIR/__init__.py:13-20 */
#define GPURIR_OMNI 0
#define GPURIR_HOMNI 1
#define GPURIR_CARD 2
#define GPURIR_HYPCARD 3
#define GPURIR_SUBCARD 4
#define GPURIR_BIDIR 5

/* Evaluation of the windowed sinc (src/gpuRIR_cuda.h:29-33 flags mixed_precision / lookup_table) */
#define GPURIR_PREC_FP32 0 /* exact FP32 sinc (reference: both flags off)            */
#define GPURIR_PREC_LUT 1  /* 1/16-sample table + linear interpolation (lookup_table) */
#define GPURIR_PREC_FP16 2 /* FP16 arithmetic, FP32 accumulation (mixed_precision): FP16 */
                           /* filter inputs on the matrix cores (basis) or packed FP16  */
                           /* windowed-sinc taps (tap scatter)                          */
#define GPURIR_PREC_BF16 3 /* BF16 filter inputs on the matrix cores, FP32 accumulation */
                           /* (basis method only; no reference counterpart: the error   */
                           /* table's BF16 entry)                                       */

/* Accumulation of the windowed sinc, in every precision mode */
#define GPURIR_METHOD_AUTO 0  /* library default: BASIS, or TAPS for launches of fewer    */
                              /* than 256 tiles of 512 early samples (too few to fill    */
                              /* the device) and for windows of more than 513 taps     */
                              /* (fs above 64 kHz)                                       */
#define GPURIR_METHOD_TAPS 1  /* every image scatters its Tw windowed-sinc taps (windows  */
                              /* of up to 3072 taps, fs up to 384 kHz)                    */
#define GPURIR_METHOD_BASIS 2 /* per-image sums on a few rows, then matrix-core FIRs:     */
                              /* FP32, 8-term Chebyshev expansion of the taps in the     */
                              /* fractional delay (<= 5e-7 per tap of 1); mixed precision */
                              /* FP16 rows and coefficients (8 terms at 16 kHz, 6 terms,  */
                              /* <= 6e-5 per tap, at other rates); table mode,            */
                              /* 16 phase rows of the 1/16-sample grid (the table's own   */
                              /* linear interpolation, exactly)                           */

/* Diffuse-tail noise */
#define GPURIR_NOISE_LOGISTIC 0 /* unit-variance logistic, as src/gpuRIR_cuda.cu:454-455 */
#define GPURIR_NOISE_GAUSSIAN 1 /* standard Gaussian (Box-Muller)                      */

typedef struct gpurir_options {
    int precision;              /* GPURIR_PREC_*                                         */
    int noise;                  /* GPURIR_NOISE_*                                        */
    unsigned long long seed;    /* Philox4x32-10 key                                     */
    unsigned int call_offset;   /* counter word 3: distinct per call                     */
    unsigned int room_index;    /* counter word 2: distinct per room                     */
    float level;                /* diffuse-tail multiplier (1 = reference)               */
    int chunk_images;           /* images per work item, 0 = planner default             */
    int tile_samples;           /* ISM samples per LDS tile, 0 = planner default         */
    int persistent;             /* batch plans: 1 = persistent accumulation launch       */
    int crossfade;              /* samples of linear cross-fade into the diffuse model   */
                                /* before Tdiff (0 = the reference's hard switch)        */
    int method;                 /* GPURIR_METHOD_*                                       */
} gpurir_options_t;

/* Library version (major*10000 + minor*100 + patch). */
int gpurir_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* gpurir_last_error(void);

/* Sample count rule nSamples = ceil(f32(T) * f32(fs)) rounded up to even:
 * src/gpuRIR_cuda.cu:831-835, src/python_bind.cpp:82-83. */
int gpurir_n_samples(float T, float fs);

/* Toggles of the drop-in state, mutually exclusive as in the reference; return the new
 * state of the flag.  Replace gpuRIR_cuda::activate_mixed_precision / activate_lut
 * (src/gpuRIR_cuda.cu:1050-1073) bound as activate_mixed_precision_bind /
 * activate_lut_bind (src/python_bind.cpp:121-128,140-142).  Mixed precision selects
 * GPURIR_PREC_FP16 for the plain entry point; its FP16 arithmetic is the packed FP16
 * and FP16 matrix-core instructions of gfx950 (the reference's Pascal-architecture check has no equivalent). */
int gpurir_activate_mixed_precision(int activate);
int gpurir_activate_lut(int activate);

/* Noise stream of the drop-in state: the reference seeds its generator with 1234
 * once (src/gpuRIR_cuda.cu:1046-1047) and the stream advances across calls; here
 * every plain gpurir_simulate_rir call uses call_offset = the number of previous calls. */
int gpurir_set_seed(unsigned long long seed, unsigned int next_call_offset);

/* simulateRIR: replaces gpuRIR_bind::simulateRIR_bind (src/python_bind.cpp:29-88)
 * -> gpuRIR_cuda::cuda_simulateRIR (src/gpuRIR_cuda.cu:766-931).  Host arrays, C order:
 * pos_src/orV_src [M_src][3], pos_rcv/orV_rcv [M_rcv][3], out [M_src][M_rcv][nSamples]
 * with nSamples = gpurir_n_samples(Tmax, fs).  Uses the drop-in precision and noise state.
 * orV_src / orV_rcv may be NULL when the corresponding pattern is omni (0). */
int gpurir_simulate_rir(const float room_sz[3], const float beta[6], const float* pos_src, int M_src,
                        const float* pos_rcv, int M_rcv, const float* orV_src, const float* orV_rcv,
                        int spkr_pattern, int mic_pattern, const int nb_img[3], float Tdiff, float Tmax,
                        float fs, float c, float* out);

/* Same with device arrays, explicit options and stream. */
int gpurir_simulate_rir_device(const float room_sz[3], const float beta[6], const float* d_pos_src, int M_src,
                               const float* d_pos_rcv, int M_rcv, const float* d_orV_src, const float* d_orV_rcv,
                               int spkr_pattern, int mic_pattern, const int nb_img[3], float Tdiff, float Tmax,
                               float fs, float c, const gpurir_options_t* opt, float* d_out, void* stream);

/* gpurir_simulate_rir_device without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results; it calls no hipMalloc,
 * hipFree or pinned-memory allocation: the launch's plan, partial rows and work counters live
 * in a static scratch array of GPURIR_NOALLOC_SCRATCH_BYTES per GPU (held until the process
 * ends), and the plan is copied from pageable host memory.  A launch that needs more scratch
 * is rejected with status -1 and a message.  Partial rows, the largest part, take 4 bytes per
 * early sample and image chunk of every RIR whose images the tap method splits over several
 * work items (about 2048 items per launch by default): early parts of up to 8000 samples fit,
 * the 1024 RIRs of 8000 samples and 246,906 images of a full trajectory do not.  Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (different streams, or threads) race and must be ordered by the caller. */
#define GPURIR_NOALLOC_SCRATCH_BYTES (64 << 20)
int gpurir_simulate_rir_device_noalloc(const float room_sz[3], const float beta[6], const float* d_pos_src,
                                       int M_src, const float* d_pos_rcv, int M_rcv, const float* d_orV_src,
                                       const float* d_orV_rcv, int spkr_pattern, int mic_pattern,
                                       const int nb_img[3], float Tdiff, float Tmax, float fs, float c,
                                       const gpurir_options_t* opt, float* d_out, void* stream);

/* Stage (a) listing: per (RIR, image) the integer sample n0 = floor(tau), the fraction
 * tau - n0 and the signed amplitude; RIR m = src*M_rcv + rcv, image i = (nx*Ny + ny)*Nz + nz.
 * Replaces calcAmpTau_kernel's amp/tau outputs (src/gpuRIR_cuda.cu:245-375). Device arrays. */
int gpurir_image_list_device(const float room_sz[3], const float beta[6], const float* d_pos_src, int M_src,
                             const float* d_pos_rcv, int M_rcv, const float* d_orV_src, const float* d_orV_rcv,
                             int spkr_pattern, int mic_pattern, const int nb_img[3], float fs, float c, int* d_n0,
                             float* d_frac, float* d_amp, void* stream);

/* Many rooms of uneven sizes in one launch (no reference counterpart: the reference
 * simulates one room per cuda_simulateRIR call, src/gpuRIR_cuda.cu:766).  Per-room
 * scalars are host arrays of n_rooms rows; positions/orientations are device arrays
 * whose rows for room k start at the sum of the previous rooms' n_src (n_rcv).  Room k
 * writes [n_src[k]][n_rcv[k]][gpurir_n_samples(Tmax[k], fs)] floats at d_out +
 * out_offsets[k] (NULL: rooms packed back to back).  level may be NULL (all 1).
 * Room k uses noise counter word room_index = opt->room_index + k. */
int gpurir_simulate_batch_device(int n_rooms, const float* room_sz, const float* beta, const int* nb_img,
                                 const int* n_src, const int* n_rcv, const float* Tdiff, const float* Tmax,
                                 const float* level, const float* d_pos_src, const float* d_pos_rcv,
                                 const float* d_orV_src, const float* d_orV_rcv, int spkr_pattern, int mic_pattern,
                                 float fs, float c, const gpurir_options_t* opt, float* d_out,
                                 const long long* out_offsets, void* stream);

/* A batch plan: the arguments of gpurir_simulate_batch_device fixed once (device
 * pointers included), uploaded once, and replayed as a captured hipGraph. */
typedef struct gpurir_plan* gpurir_plan_t;
int gpurir_batch_plan_create(int n_rooms, const float* room_sz, const float* beta, const int* nb_img,
                             const int* n_src, const int* n_rcv, const float* Tdiff, const float* Tmax,
                             const float* level, const float* d_pos_src, const float* d_pos_rcv,
                             const float* d_orV_src, const float* d_orV_rcv, int spkr_pattern, int mic_pattern,
                             float fs, float c, const gpurir_options_t* opt, float* d_out,
                             const long long* out_offsets, gpurir_plan_t* plan);
/* Enqueue the plan on a stream (graph launch after the first capture). */
int gpurir_batch_plan_run(gpurir_plan_t plan, void* stream);
/* Total output floats of the plan and number of RIRs. */
long long gpurir_batch_plan_out_elems(gpurir_plan_t plan);
long long gpurir_batch_plan_n_rirs(gpurir_plan_t plan);
int gpurir_batch_plan_destroy(gpurir_plan_t plan);
/* Reflection coefficients and diffuse levels read from device memory at every run (no
 * reference counterpart: the reference takes beta per cuda_simulateRIR call,
 * src/gpuRIR_cuda.cu:766, and has no level).  d_beta [n_rooms][6] and d_level [n_rooms]
 * float32 (either may be NULL: the values given at creation) must stay allocated while
 * the plan runs; their contents may change between runs without re-planning, so an
 * optimiser can update betas, levels and (in the position arrays given at creation)
 * sources and receivers in place and replay the same graph.  Applies to the plan's runs,
 * autotuning, profiling and gpurir_batch_plan_adjoint.  Invalidates the captured graph. */
int gpurir_batch_plan_bind_params(gpurir_plan_t plan, const float* d_beta, const float* d_level);
/* Autotuner (no reference counterpart): times the plan's launch eagerly on the stream
 * for candidate work-item shapes -- ISM samples per LDS tile, images per work item --
 * and for the persistent launch form, keeps the fastest and re-plans with it (the next
 * run captures a new graph).  Choices are cached per process by GPU architecture,
 * precision and workload signature.  Synchronous; overwrites the plan's output.
 * Outputs (any may be NULL): the chosen images per item and samples per tile (0 =
 * planner default), 1 if the persistent form was chosen, and its time in ms. */
int gpurir_batch_plan_autotune(gpurir_plan_t plan, void* stream, int* chunk_images, int* tile_samples,
                               int* persistent, float* ms);
/* Algorithmic work of the plan: the number of (RIR, image, t) with 0 <= t < nISM and
 * |t - tau| < Tw/2 (Tw/2 rounded down in table mode).  Synchronous.  No reference
 * counterpart (measurement support). */
int gpurir_batch_plan_count_taps(gpurir_plan_t plan, long long* taps, void* stream);
/* Runs the plan's two kernels eagerly on the stream and returns the duration of each,
 * measured with HIP events recorded on that stream.  Synchronous.  No reference
 * counterpart (measurement support). */
int gpurir_batch_plan_profile(gpurir_plan_t plan, void* stream, float* ms_accumulate, float* ms_finalize);

/* Adjoint of gpurir_simulate_batch_device (no reference counterpart: gpuRIR is not
 * differentiable).  Given the forward output d_out (same arguments, options and noise
 * stream) and d_grad_out = dL/d out (same layout), writes dL/d beta [n_rooms][6],
 * dL/d pos_src and dL/d pos_rcv (rows as the position arrays) and dL/d level
 * [n_rooms]; any gradient pointer may be NULL.  The windowed sinc and its derivative
 * are evaluated exactly (FP32 taps, FP64 chain rule) whatever opt->precision is; the
 * integer envelope window start is held fixed.  Deterministic. */
int gpurir_simulate_batch_adjoint_device(int n_rooms, const float* room_sz, const float* beta, const int* nb_img,
                                         const int* n_src, const int* n_rcv, const float* Tdiff, const float* Tmax,
                                         const float* level, const float* d_pos_src, const float* d_pos_rcv,
                                         const float* d_orV_src, const float* d_orV_rcv, int spkr_pattern,
                                         int mic_pattern, float fs, float c, const gpurir_options_t* opt,
                                         const float* d_out, const float* d_grad_out, const long long* out_offsets,
                                         float* d_grad_beta, float* d_grad_src, float* d_grad_rcv,
                                         float* d_grad_level, void* stream);
/* The same for a batch plan, using the plan's output buffer as the forward output. */
int gpurir_batch_plan_adjoint(gpurir_plan_t plan, const float* d_grad_out, float* d_grad_beta, float* d_grad_src,
                              float* d_grad_rcv, float* d_grad_level, void* stream);

/* gpu_conv: replaces gpuRIR_bind::gpu_conv (src/python_bind.cpp:90-119) ->
 * gpuRIR_cuda::cuda_convolutions (src/gpuRIR_cuda.cu:933-1023).  Host arrays:
 * segments [M_src][seg_len], rirs [M_src][M_rcv][rir_len],
 * out [M_src][M_rcv][seg_len + rir_len - 1] (linear convolutions, of any length: beyond
 * 32768 samples the segment and the RIR are cut into blocks whose products are added). */
int gpurir_conv(const float* segments, int M_src, int seg_len, const float* rirs, int M_rcv, int rir_len,
                float* out);

/* Whole simulateTrajectory (gpuRIR/__init__.py:177-221: segmentation, gpu_conv and the
 * float64 overlap-add) on the device.  d_signal [n_signal] float32, d_rirs
 * [n_pts][n_rcv][rir_len] float32, w_ini host [n_pts + 1] segment starts (w_ini[n_pts]
 * = n_signal, non-decreasing), d_out [n_signal + rir_len - 1][n_rcv] float64. */
int gpurir_trajectory_device(const float* d_signal, int n_signal, const float* d_rirs, int n_pts, int n_rcv,
                             int rir_len, const int* w_ini, double* d_out, void* stream);

/* A trajectory plan: gpurir_trajectory_device bound to fixed device buffers (d_signal,
 * d_rirs, d_out, as above) and one segmentation w_ini (host, copied).  The plan owns its
 * tables and workspace; each run enqueues the simulateTrajectory of the buffers' current
 * contents on `stream`.  Segments of S or S + 1 samples (256 <= S <= 1536, the segmentation
 * of simulateTrajectory without timestamps, gpuRIR/__init__.py:204-208) launch the plan's
 * three kernels directly (no host work per run; a caller may capture them in its own graph);
 * any other segmentation runs gpurir_trajectory_device eagerly.
 * No counterpart in the reference: it serves device-resident pipelines that call
 * simulateTrajectory repeatedly on the same shapes. */
typedef struct gpurir_traj_plan* gpurir_traj_plan_t;
int gpurir_trajectory_plan_create(const float* d_signal, int n_signal, const float* d_rirs, int n_pts, int n_rcv,
                                  int rir_len, const int* w_ini, double* d_out, gpurir_traj_plan_t* plan);
int gpurir_trajectory_plan_run(gpurir_traj_plan_t plan, void* stream);
int gpurir_trajectory_plan_destroy(gpurir_traj_plan_t plan);

#ifdef __cplusplus
}
#endif
#endif /* GPURIR_H */
