/*
 * nf3d.h -- C ABI of the MI355X (gfx950) neo_fft3d denoiser.
 *
 * Drop-in boundary.  In the reference the per-plane engine FFT3DEngine
 * (src/fft3d_engine.h:22-741) and the FFTW library behind it
 * (src/fftwlite.h:21-96) sit under the FFT3D front end (src/fft3d.hpp:12-364),
 * which the DualSynth bridge exposes to AviSynth+ as neo_fft3d() and to
 * VapourSynth as neo_fft3d.FFT3D (include/dualsynth/avs_wrapper.hpp:211-228,
 * vs_wrapper.hpp:195-212).  This library replaces the engine and FFTW with HIP
 * kernels and also offers the front end itself, so a plugin binds either level:
 *
 *   engine level   nf3d_engine_*   one engine per processed plane, exactly
 *                                  where FFT3DEngine is constructed and called
 *   front-end      nf3d_filter_*   parameter parsing, defaults, plane
 *                                  selection, crop copy, frame-request window
 *
 * Conventions: functions return 0 on success and -1 on failure with a
 * NUL-terminated message in err (the reference's own wording for the errors
 * it raises; the caller prefixes "FFT3D: " / "neo_fft3d: " as the bridge
 * does).  Host pointers are borrowed for the duration of a call; device
 * memory is owned by the library.  Plain C types only.
 *
 * Streams: calls that take a stream enqueue their work on it (NULL: the
 * engine's own non-blocking stream, which does not wait for work the caller
 * queued on other streams -- the caller orders its producers before the
 * call).  Successive calls on one engine are ordered by the library: when a
 * call uses another stream than the engine's previous call, its stream first
 * waits for that call's work (the spectra ring, Kalman state and seam buffers
 * carry over between calls).  Graph replays count as calls on the replay
 * stream.
 */
#ifndef NF3D_H
#define NF3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Every knob of EngineParams (src/fft3d_common.h:48-77), as the user gave it
 * (bit-depth scaling of sigma/smin/smax happens inside, as in
 * src/fft3d_engine.h:91-103). */
typedef struct nf3d_params {
  float sigma, beta;
  int bw, bh, bt, ow, oh;
  float kratio, sharpen, scutoff, svr, smin, smax;
  int measure, interlaced, wintype, pframe, px, py, pshow;
  float pcutoff, pfactor, sigma2, sigma3, sigma4, degrid, dehalo, hr, ht;
  int l, t, r, b, opt;
} nf3d_params;

/* Clip format, as DSFormat (include/dualsynth/ds_format.hpp:10-26). */
typedef struct nf3d_format {
  int color_family;    /* 0 gray, 1 yuv, 2 rgb */
  int is_float;        /* 32-bit float samples */
  int bits_per_sample; /* 8, 9..16 (uint16 storage), 32 */
  int ssw, ssh;        /* chroma subsampling shifts (yuv) */
  int num_planes;      /* 1 (gray), 3, or 4: YUVA / RGBA, the alpha plane at
                        * the luma size; copied unless the `planes` list
                        * names it (the AviSynth+ y/u/v modes never do,
                        * src/fft3d.hpp:149-169) */
} nf3d_format;

/* One plane of one frame.  `data` is host memory for the *_get_frame calls
 * and device memory for the *_device calls; stride in bytes. */
typedef struct nf3d_frame_ref {
  int n;
  const void* data;
  int stride;
} nf3d_frame_ref;

typedef struct nf3d_engine nf3d_engine;

/* Front-end defaults (src/fft3d.hpp:77-106): sigma2..sigma4 = sigma. */
void nf3d_params_defaults(nf3d_params* p, float sigma);

/* Replaces `new FFT3DEngine(ep, plane, ...)` (src/fft3d.hpp:186-198,
 * src/fft3d_engine.h:87-407): validation, geometry, windows and tables, and
 * the engine's device state on the current HIP device.  width/height are the
 * clip's (luma) size; plane selects luma or chroma geometry. */
int nf3d_engine_create(const nf3d_params* p, const nf3d_format* fmt, int width, int height,
                       int frames, int plane, nf3d_engine** out, char* err, size_t errlen);

/* Replaces FFT3DEngine::GetFrame(n, in_frames) (src/fft3d_engine.h:452-734).
 * `frames` must hold this plane of every frame in the request window
 * [max(n - bt/2, 0), min(n + (bt-1)/2, frames-1)] and, while the noise
 * pattern is still to be estimated (pfactor != 0), of frame pframe.
 * The filtered plane is written into dst's processed region (the plane minus
 * the l/t/r/b crop); crop borders are left untouched.  Thread-safe: calls on
 * one engine are serialised (bt == 0 is order dependent, as in the reference). */
int nf3d_engine_get_frame(nf3d_engine* e, int n, const nf3d_frame_ref* frames, int nframes,
                          void* dst, int dst_stride, char* err, size_t errlen);

/* As nf3d_engine_get_frame, with device-resident planes: the work is
 * enqueued on `stream` (a hipStream_t, or NULL for the engine's own stream)
 * and the call returns without waiting.  dst may not alias a source plane. */
int nf3d_engine_get_frame_device(nf3d_engine* e, int n, const nf3d_frame_ref* frames, int nframes,
                                 void* dst, int dst_stride, void* stream, char* err, size_t errlen);

/* Throughput form of nf3d_engine_get_frame_device for a device-resident clip
 * (no reference counterpart; the reference is driven one GetFrame at a time):
 * filters frames [first, first+count) in order.  src[k] / dst[k] are this
 * plane of frame k of the clip (k < frames), all with the given strides.  The
 * work is enqueued on `stream`; the call returns without waiting. */
int nf3d_engine_run_device(nf3d_engine* e, int first, int count, const void* const* src, int src_stride,
                           void* const* dst, int dst_stride, void* stream, char* err, size_t errlen);

/* Batched form of nf3d_engine_run_device (no reference counterpart): one
 * item per engine -- the planes of a clip, or of several clips -- each with
 * its own frame range and planes.  The result is that of calling
 * nf3d_engine_run_device for every item, bit for bit; engines with the same
 * block shape, mode (bt) and sharpen/dehalo case share their frame-loop
 * launches (up to 6 engines per launch), on `stream` (NULL: the first
 * item's engine's stream); groups of different shapes may run side by side
 * on their first engines' own streams, forked from `stream` and joined back
 * to it, so the call's work ends on `stream` (one wait there covers every
 * item).  An engine may appear once per call. */
typedef struct {
  nf3d_engine* engine;
  int first, count;
  const void* const* src;
  int src_stride;
  void* const* dst;
  int dst_stride;
} nf3d_run_item;
int nf3d_run_device_batch(const nf3d_run_item* items, int n, void* stream, char* err, size_t errlen);

/* nf3d_run_device_batch with the frames in host memory (no reference
 * counterpart: the host-fed form of FFT3D::GetFrame driven over a range,
 * src/fft3d.hpp:201-264).  Every item covers the same frame range
 * [first, first+count) of clips of the same length; src[k] / dst[k] are host
 * planes of frame k (the source for every frame the filter windows read,
 * up to two beyond each end of the range).  The frames stream through the
 * device in chunks: the caller's planes are copied into pinned staging by
 * host threads and uploaded on one stream while the previous chunk is
 * filtered on another and the one before is downloaded on a third.  Only the
 * processed region (the plane minus the l/t/r/b crop) of dst is written.
 * Returns when every frame is in dst. */
int nf3d_run_host_batch(const nf3d_run_item* items, int n, char* err, size_t errlen);

/* The work of one nf3d_run_device_batch call captured as a HIP graph on
 * `stream` (a non-null hipStream_t) instead of being run; nf3d_graph_launch
 * replays it on a stream (each replay filters the same frames again, as a
 * repeated batch call would).  Capture needs the engines' pattern estimate
 * already made (a first, uncaptured call does it) since it synchronises. */
typedef struct nf3d_graph nf3d_graph;
int nf3d_run_device_batch_graph(const nf3d_run_item* items, int n, void* stream, nf3d_graph** graph, char* err,
                                size_t errlen);
/* The same capture bound to the engines' spectra rings as they stand (no
 * reference counterpart; for a stream of frame steps replayed in order): the
 * captured work reads the Wiener-3D spectra the rings hold instead of
 * transforming those frames again.  nf3d_graph_launch then replays it only
 * if every engine has moved on by a whole number of frames since the
 * capture (its latest call's last frame, a multiple of bt) and its ring holds
 * the frames it held at capture shifted by that many (frame f sits in slot
 * f mod bt), and a learned PSD kept from a span before the capture is the
 * one of the span shifted alike (the shift a multiple of the period);
 * otherwise it returns -1 without launching.  The caller puts the shifted
 * frames at the captured frame pointers (staging slots reused periodically,
 * as a serving loop does). */
int nf3d_run_device_batch_graph_ring(const nf3d_run_item* items, int n, void* stream, nf3d_graph** graph,
                                     char* err, size_t errlen);
int nf3d_graph_launch(nf3d_graph* graph, void* stream);
void nf3d_graph_destroy(nf3d_graph* graph);

/* DFT-as-GEMM precision form (no reference counterpart): unnormalised 2D
 * DFTs of n x n real blocks (n = 32, 48 or 64) as dense DFT-matrix products on the
 * matrix cores, FP16 hi/lo split operands (hi*hi + hi*lo + lo*hi) with FP32
 * accumulation.  Forward: in [nblocks][n][n] float -> out [nblocks][n][n/2+1]
 * complex (FFTW r2c layout); inverse: in [nblocks][n][n/2+1] complex -> out
 * [nblocks][n][n] float (FFTW c2r: the imaginary parts of columns 0 and n/2
 * are ignored).  `scale`, a power of two, multiplies the input before it is
 * split so every operand stays in FP16 range (|scale * x| < 32768 forward,
 * < 65504 inverse); the output is corrected exactly.  Device pointers; the
 * work is enqueued on `stream`.  0 on success. */
int nf3d_dft2d_mfma(const void* in, void* out, int nblocks, int n, int inverse, float scale, void* stream);

/* ------------------------------------------------------------ learned PSD
 * No reference counterpart.  A small MLP maps a block's log-power spectrum
 * (n_in = bh * (bw/2+1) values) through two GELU hidden layers to a per-block
 * noise log-PSD (n_out values); the engine can use it in place of the single
 * noise pattern of src/helper.cpp:57-132 (per-block pattern in the Wiener
 * gains).  BF16 matrix-core kernels with FP32 accumulation; FP32 master
 * weights and Adam state.  Parameters in host order W1 [n_in][h], b1 [h],
 * W2 [h][h], b2 [h], W3 [h][n_out], b3 [n_out].  Features and batches are
 * BF16 device arrays, rows of n_in.  0 on success, -1 with `err` set. */
typedef struct nf3d_mlp nf3d_mlp;
int nf3d_mlp_create(int n_in, int n_hidden, int n_out, int max_batch, nf3d_mlp** out, char* err, size_t errlen);
void nf3d_mlp_destroy(nf3d_mlp* m);
long long nf3d_mlp_param_count(const nf3d_mlp* m);
void nf3d_mlp_dims(const nf3d_mlp* m, int* out3);  /* n_in, n_hidden, n_out */
int nf3d_mlp_set_params(nf3d_mlp* m, const float* host, char* err, size_t errlen);  /* resets Adam */
int nf3d_mlp_get_params(nf3d_mlp* m, float* host, char* err, size_t errlen);
int nf3d_mlp_get_grads(nf3d_mlp* m, float* host, char* err, size_t errlen);
/* Feature x_k = (ln(|c_k|^2 + eps) - mu_k) * isd_k of a compact spectrum c. */
int nf3d_mlp_set_input_norm(nf3d_mlp* m, const float* mu, const float* isd, float eps, char* err, size_t errlen);
int nf3d_mlp_features(nf3d_mlp* m, const void* spec, int n, void* x_out, void* stream);
/* FP32 -> BF16 (round to nearest even); |c|^2 of `count` complex values. */
int nf3d_to_bf16(const float* x, void* out, long long count, void* stream);
int nf3d_periodogram(const void* spec, float* out, long long count, void* stream);
/* Training batches: rows idx[0..rows) of src into dst (row_bytes a multiple
 * of 16; device pointers, int32 row indices). */
int nf3d_gather_rows(const void* src, const int* idx, void* dst, int rows, int row_bytes, void* stream);
/* Inference over n rows: out = log-PSD, or exp(log-PSD) * scale when
 * exp_out; `work` holds nf3d_mlp_work_bytes(m, n) bytes. */
long long nf3d_mlp_work_bytes(const nf3d_mlp* m, int n);
int nf3d_mlp_predict(nf3d_mlp* m, const void* x, int n, float* out, int exp_out, float scale, void* work,
                     void* stream, char* err, size_t errlen);
/* nf3d_mlp_predict without a workspace and without any allocation: the
 * hidden activations live in a fixed array the library holds on each GPU
 * (16 MiB: n * n_hidden <= 4,194,304, i.e. 16384 rows at 256 hidden units;
 * larger n is refused with an error before any launch).  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, hosts that forbid hipMalloc).  Every
 * call on one GPU shares that GPU's array, so calls that overlap on one GPU
 * (different streams or threads without ordering) race; each call writes
 * every element it reads, so nothing carries over between calls. */
int nf3d_mlp_predict_noalloc(nf3d_mlp* m, const void* x, int n, float* out, int exp_out, float scale, void* stream,
                             char* err, size_t errlen);
/* Gradients of the Whittle loss mean(y + t e^-y) (y = predicted log-PSD,
 * t = periodogram targets [n][n_out] FP32) over a batch of 1..max_batch rows;
 * `loss` (optional) waits for the stream.  train_step adds an Adam update
 * (beta 0.9 / 0.999, eps 1e-8). */
int nf3d_mlp_grad(nf3d_mlp* m, const void* x, const float* t, int n, double* loss, void* stream, char* err,
                  size_t errlen);
int nf3d_mlp_adam(nf3d_mlp* m, float lr, float beta1, float beta2, float eps, void* stream, char* err,
                  size_t errlen);
int nf3d_mlp_train_step(nf3d_mlp* m, const void* x, const float* t, int n, float lr, double* loss, void* stream,
                        char* err, size_t errlen);
/* The plane's own analysis (FrameToCover + window + r2c, as the filter
 * computes it) of one device plane into compact spectra [nb][bh][bw/2+1]. */
int nf3d_engine_spectra(nf3d_engine* e, const void* src, int src_stride, void* spec_out, void* stream, char* err,
                        size_t errlen);
/* The per-region noise-pattern estimate of one device plane as a per-block
 * PSD (no reference counterpart; the estimator of
 * nf3d_engine_set_pattern_regions applied to `src`, whatever pfactor is):
 * out[nb][bh][bw/2+1] (device, FP32) holds, for every block, its rx x ry
 * region's pattern |X - degrid correction|^2 x pattern window
 * (src/helper.cpp:92-132).  A second label source for the learned PSD: it
 * needs no knowledge of the noise.  The engine's own pattern is untouched. */
int nf3d_engine_region_psd(nf3d_engine* e, const void* src, int src_stride, int rx, int ry, float* out,
                           void* stream, char* err, size_t errlen);
/* Use the model's per-block PSD (times `strength`) as the noise pattern of
 * every frame (null model: back to the reference's sigma / pattern).  The
 * model must be n_in = n_out = bh * (bw/2+1) and must not be trained while
 * an engine uses it.  The per-frame path predicts each frame's PSD from that
 * frame; the frame-loop path (nf3d_engine_run_device, the batch calls) predicts
 * it once per launch, from the launch's first frame, for all the launch's
 * frames (frames per launch: nf3d_engine_set_tuning; 1 gives the per-frame
 * semantics). */
int nf3d_engine_set_learned(nf3d_engine* e, nf3d_mlp* m, float strength, char* err, size_t errlen);
/* Prediction period of the learned PSD in the frame-loop path (no reference
 * counterpart; the reference's own pattern is estimated once, from frame
 * pframe, for the whole clip): 0 (default) = one prediction per launch, as
 * above; n > 0 = one per span of frames [k n, (k + 1) n), made from the
 * span's first frame-loop frame that a call reaches and kept for the calls
 * that continue the span (in-order calls of any length give the PSD of
 * frame max(k n, first frame-loop frame) to every frame-loop frame of span
 * k).  Launches are split at the span starts, so the result no longer
 * depends on frames per launch.  A per-frame call, a new model or period,
 * or a graph capture not bound to the rings (which predicts at its first
 * launch) starts over.  Errors: n < 0. */
int nf3d_engine_set_learned_period(nf3d_engine* e, int frames, char* err, size_t errlen);

/* Spatially varying noise pattern (no reference counterpart; it generalises
 * the pattern estimate of src/helper.cpp:46-132): with pfactor != 0 the
 * block grid is cut into regions of rx x ry blocks, each region takes the
 * pattern of its own first strict minimum of FindPatternBlock's sums over
 * the search range's blocks inside it (or the plane's choice when it has
 * none), and every block is filtered with its region's pattern (times bt in
 * the 3D gains).  rx = ry = 0 restores the single pattern.  Takes effect at
 * the next estimate (the next call, on frame pframe).  Errors: no estimated
 * pattern (pfactor == 0, pattern display, sigma2..sigma4), sizes. */
int nf3d_engine_set_pattern_regions(nf3d_engine* e, int rx, int ry, char* err, size_t errlen);

/* Precision form of the engine's 2D transforms (no reference counterpart; the
 * reference uses FFTW): mode 0 = FP32 FFT in registers (default), 1 = dense
 * DFT-matrix products on the matrix cores with FP16 hi/lo split operands and
 * FP32 accumulation (32x32, 48x48 and 64x64 blocks, integer formats: the FP16
 * operand scales come from the format's sample range).  Errors: other block
 * shapes, float formats, other modes. */
int nf3d_engine_set_dft(nf3d_engine* e, int mode, char* err, size_t errlen);

/* ------------------------------------------------------------ launch tuning
 * No reference counterpart (the reference's `opt` picks an ISA; these pick
 * launch shapes, and no choice changes a result).  The frame-loop path runs
 * frames_per_launch frames per kernel launch (1..16) and, for Wiener 2D and
 * sharpen-only (no temporal recursion), frames_per_workgroup frames per
 * workgroup (Wiener 3D and Kalman keep a launch's frames in one workgroup).
 * A process-wide table, one line per "bw bh bits bt frames_per_launch
 * frames_per_workgroup" ('#' comments; bits 32 = float), is read from the file
 * named by NF3D_TUNE_TABLE or by nf3d_tune_load (which replaces it and returns
 * its line count); engines created afterwards take their entry, except that
 * attaching a learned PSD (whose prediction is made once per launch, so that
 * frames per launch is part of its result, unless a prediction period is
 * set) restores the default 16 frames per launch unless
 * nf3d_engine_set_tuning set it.
 * neo_fft3d_amd/autotune.py measures the choices and writes such a table. */
int nf3d_tune_load(const char* path, char* err, size_t errlen);
int nf3d_tune_lookup(int bw, int bh, int bits, int bt, int* out2);  /* 1 if found */
int nf3d_engine_set_tuning(nf3d_engine* e, int frames_per_launch, int frames_per_workgroup, char* err,
                           size_t errlen);
int nf3d_engine_get_tuning(nf3d_engine* e, int* out2);

/* Replaces FFT3DEngine::CacheRefresh (src/fft3d_engine.h:736-740): 1 if the
 * spectrum of frame n is held in the HBM spectra ring (and marks it recent). */
int nf3d_engine_cache_refresh(nf3d_engine* e, int n);

/* Block geometry of the engine: nox, noy, outwidth, mirw, mirh, ow, oh. */
int nf3d_engine_geometry(const nf3d_engine* e, int* out7);

/* Diagnostics (no reference counterpart): when enabled, every filtering
 * launch (the spectral kernel with its seam pass; one frame, or a run of
 * frames in nf3d_engine_run_device) is bracketed by HIP events on its stream;
 * nf3d_engine_profile_read waits for them and returns the summed kernel time
 * (ms) and the number of frames those launches filtered since the last read. */
int nf3d_engine_profile(nf3d_engine* e, int enable);
int nf3d_engine_profile_read(nf3d_engine* e, double* total_ms, int* frames);

/* Wait for the engine's queued work (the host-fed calls already do). */
int nf3d_engine_sync(nf3d_engine* e);

void nf3d_engine_destroy(nf3d_engine* e);

/* --------------------------------------------------------------- front end
 * The FFT3D filter of src/fft3d.hpp: 41 parameters with the reference's
 * defaults, `planes` (VapourSynth) or y/u/v (AviSynth+) plane modes, l/t/r/b
 * crop with border copy, and the frame-request window. */
typedef struct nf3d_filter nf3d_filter;

/* Argument signatures generated by the bridge from FFT3D::Params()
 * (include/dualsynth/ds_filter.hpp:19-69, src/fft3d.hpp:29-73). */
const char* nf3d_vs_signature(void);
const char* nf3d_avs_signature(void);

/* Replaces FFT3D::Initialize (src/fft3d.hpp:74-199).  `args` is the
 * invocation's arguments as "name=value;name=value" (integer lists
 * comma-separated, e.g. "planes=0,2"); absent names take the defaults.
 * Passing y/u/v instead of planes selects the AviSynth+ plane modes. */
int nf3d_filter_create(const char* args, const nf3d_format* fmt, int width, int height, int frames,
                       nf3d_filter** out, char* err, size_t errlen);

/* FFT3D::RequestReferenceFrames (src/fft3d.hpp:201-219), plus frame pframe
 * while the noise pattern is pending.  Returns the count (<= cap). */
int nf3d_filter_request_frames(nf3d_filter* f, int n, int* out, int cap);

/* 0: serial (bt == 0, fmSerial / MT_SERIALIZED), 1: parallel
 * (fmParallel / MT_NICE_FILTER) -- src/fft3d.hpp:27-28. */
int nf3d_filter_mt_mode(const nf3d_filter* f);

/* One frame of the clip: all planes, host memory. */
typedef struct nf3d_frame_planes {
  int n;
  const void* data[4];
  int stride[4];
} nf3d_frame_planes;

/* FFT3D::GetFrame (src/fft3d.hpp:221-264): `frames` holds the frames named by
 * nf3d_filter_request_frames(n); every plane of frame n is written to dst. */
int nf3d_filter_get_frame(nf3d_filter* f, int n, const nf3d_frame_planes* frames, int nframes,
                          void* const* dst, const int* dst_stride, char* err, size_t errlen);

void nf3d_filter_destroy(nf3d_filter* f);

/* ------------------------------------------------------------ clip-level
 * A whole in-memory clip behind the front end, with the same signature as the
 * oracle and reference test drivers (family 0 gray / 1 yuv / 2 rgb, 3 yuva /
 * 4 rgba with a fourth plane, sample_type 0 integer / 1 float,
 * planes[frame * nplanes + plane]). */
void* nf3d_clip_create(const char* args, int family, int sample_type, int bits, int ssw, int ssh,
                       int width, int height, int nframes, const void* const* planes,
                       const int* strides, char* err, int errlen);
int nf3d_clip_get_frame(void* h, int n, void* const* dst, const int* dst_strides, char* err,
                        int errlen);
void nf3d_clip_free(void* h);

/* Library identity: "Neo FFT3D Filter (MI355X/gfx950) <version>". */
const char* nf3d_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NF3D_H */
