/*
 * npx_esrrf.h — C ABI of the MI355X (gfx950) eSRRF library (libnpx_hip.so).
 *
 * The drop-in boundary for NanoPyx's eSRRF hot path: the Python host side (nanopyx_amd, the new
 * "HIP_gfx950" run type of the eSRRF and RadialGradientConvergence Liquid Engine classes) binds these
 * with ctypes. Plain pointers and sizes only: every buffer is a caller-owned device pointer, `stream` is a
 * hipStream_t (NULL = the default stream), and calls are asynchronous on that stream. Calls on distinct
 * streams are reentrant; the library's only shared state is a mutex-protected cache of per-(size,
 * magnification, radius) geometry tables. Status: 0 = OK, < 0 = error, message in npx_last_error()
 * (thread-local). Paths in citations are relative to the NanoPyx repository root (v2.1.2).
 */
#ifndef NPX_ESRRF_H
#define NPX_ESRRF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NPX_OK 0
#define NPX_ERR_ARG (-1)         /* bad shape / parameter (reference: ValueError / TypeError)        */
#define NPX_ERR_HIP (-2)         /* HIP runtime error (launch, allocation, copy)                     */
#define NPX_ERR_UNSUPPORTED (-3) /* valid for the reference but outside what these kernels build      */
#define NPX_ERR_WORKSPACE (-4)   /* workspace pointer given but smaller than required                 */
#define NPX_ERR_INTERNAL (-5)    /* an internal consistency check failed (a library bug)              */

#define NPX_DTYPE_F32 0
#define NPX_DTYPE_U16 1 /* uint16 camera counts, converted in-kernel exactly like astype(float32) */

#define NPX_AVG 1u
#define NPX_VAR 2u
#define NPX_TAC2 4u

/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int npx_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *npx_last_error(void);

/* Writes "<device name>;<gcnArchName>;<CUs>;<clock kHz>" of the current device into buf. */
int npx_device_info(char *buf, int len);

/*
 * Per-frame eSRRF maps: d_out (N, H*M, W*M) float32 with a zero border of int(float(R)*M) pixels.
 * Replaces eSRRF._run_threaded / _run_opencl, src/nanopyx/core/transform/_le_esrrf.pyx:172-195 and :39-170
 * (CR x M, Roberts cross, CR x 2M of both gradients, RGC x intensity), which the reference binds through
 * Cython (eSRRF.run, _le_esrrf.pyx:30-33). `pad_edges` has no counterpart: the reference does not forward
 * it on this path (:194). d_in is (N, H, W) of `dtype`.
 */
int npx_esrrf_frames(const void *d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                     float *d_out, void *stream);

/* Bytes of device workspace npx_esrrf_temporal uses to split the frames of one call over more blocks. */
size_t npx_esrrf_temporal_workspace(int N, int H, int W, int M, double R);

/*
 * Fused eSRRF + temporal correlation: for each bit of mode_mask writes one (H*M, W*M) map — AVG, VAR
 * (ddof 0) and/or TAC2 — of the per-frame eSRRF maps without materialising them. Replaces the composition
 * eSRRF_ST.run -> calculate_eSRRF_temporal_correlations of src/nanopyx/methods/esrrf/eSRRF_workflow.py:98-119
 * (temporal step src/nanopyx/core/transform/sr_temporal_correlations.py:46-75,205-221). d_work may be NULL
 * (work_bytes 0): the frames are then not split. Output pointers of unset mode bits may be NULL.
 */
int npx_esrrf_temporal(const void *d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                       unsigned mode_mask, float *d_avg, float *d_var, float *d_tac2, void *d_work, size_t work_bytes,
                       void *stream);

/*
 * npx_esrrf_temporal without any device allocation, for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. The geometry tables and
 * the frame-chunk partial moments live in a static device array of 64 MiB (one per GPU, held until the process
 * ends); the tables are copied into it on `stream` by each call (hipMemcpyAsync from host memory). The frames are
 * split as npx_esrrf_temporal splits them when its partials fit beside the tables, otherwise into as many chunks as
 * fit (maps then differ from npx_esrrf_temporal's only in the float32 rounding of the chunk merge). Image sizes
 * whose tables exceed the array are refused (NPX_ERR_ARG). Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams, other threads) race and must be serialised by the caller.
 */
int npx_esrrf_temporal_static(const void *d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                              unsigned mode_mask, float *d_avg, float *d_var, float *d_tac2, void *stream);

/*
 * RadialGradientConvergence on caller arrays: d_gx, d_gy (N, 2*HM, 2*WM), d_img (N, HM, WM), indexed exactly as
 * _c_calculate_rgc does (src/include/_c_sr_radial_gradient_convergence.c:17-90). When HM and WM are multiples of M
 * (and the radius is within the eSRRF kernels' range) the fused eSRRF kernel runs on the lattice the reference reads
 * (every M-th position of the x2M arrays); other shapes run a per-pixel kernel. Replaces
 * RadialGradientConvergence._run_threaded / _run_opencl, src/nanopyx/core/transform/_le_radial_gradient_convergence.pyx:79-115
 * and :229-332 (pad_edges -> margin 0, :92-95).
 */
int npx_rgc(const float *d_gx, const float *d_gy, const float *d_img, int N, int HM, int WM, int M, double R,
            double S, int iw, int pad_edges, float *d_out, void *stream);
/*
 * npx_rgc with the gradients given only at the lattice it reads, compact: d_gx, d_gy (N, 2H + 1, 2W + 1) with
 * H = HM / M, W = WM / M, entry (i, j) = the x2M array's value at row min(i M, 2 HM - 1), column min(j M, 2 WM - 1)
 * (npx_fht_interpolate_lattice with stride M writes this layout). Maps bit-identical to npx_rgc on the full arrays.
 * Runs on the fused kernel only: NPX_ERR_UNSUPPORTED when M < 2, HM or WM is not a multiple of M, M > 64 or the
 * radius is outside the eSRRF kernels' range. For eSRRF's use_fht branch (_le_esrrf.pyx:183-188).
 */
int npx_rgc_lattice(const float *d_gx, const float *d_gy, const float *d_img, int N, int HM, int WM, int M,
                    double R, double S, int iw, int pad_edges, float *d_out, void *stream);

/*
 * AVG / VAR / TAC2 over axis 0 of a device (N, P) float32 stack. Replaces calculate_eSRRF_temporal_correlations,
 * src/nanopyx/core/transform/sr_temporal_correlations.py:46-75 (and calculate_tac2, :205-221).
 */
int npx_temporal(const float *d_stack, int N, long long P, unsigned mode_mask, float *d_avg, float *d_var,
                 float *d_tac2, void *stream);

/* Bytes of device workspace npx_esrrf_sweep uses to split the frames of one sweep over more blocks. */
size_t npx_esrrf_sweep_workspace(int N, int H, int W, int n_mag, const int *mags, int n_rad, const double *radii,
                                 int n_sens);

/*
 * Parameter sweep: the temporal AVG map of eSRRF for every (magnification, radius, sensitivity) of the three lists,
 * from one device stack, in one batched launch (plus one merge launch per magnification when d_work is given).
 * Sensitivities share the ring sums (only the final power differs) and the radii of one magnification share the
 * gradient lattice. d_avg is a host array of n_mag * n_rad * n_sens device pointers; map (m, r, s) is
 * d_avg[(m * n_rad + r) * n_sens + s], (H*mags[m], W*mags[m]) float32. Replaces the loop of
 * ParameterSweep.run, src/nanopyx/core/analysis/parameter_sweep.py:33-42 (one eSRRF(...).run per (sensitivity,
 * radius) at fixed magnification, :35-37, then calculate_eSRRF_temporal_correlations(rgc_map, "AVG"), :40).
 * d_work may be NULL (work_bytes 0): the frames are then not split.
 */
int npx_esrrf_sweep(const void *d_in, int dtype, int N, int H, int W, int n_mag, const int *mags, int n_rad,
                    const double *radii, int n_sens, const double *sens, int iw, float *const *d_avg, void *d_work,
                    size_t work_bytes, void *stream);

/*
 * HIP graph capture of a launch sequence. npx_capture_begin creates a private stream in capture mode; every npx_*
 * call made on it until npx_capture_end is recorded (not executed), and npx_capture_end returns an instantiated
 * graph that npx_graph_launch replays on any stream. Calls made under capture need their geometry tables on the
 * device already (make one ordinary call with the same parameters first); they record no memset nodes. No reference
 * counterpart: it replaces the reference's per-call launch sequence (.wait() after each OpenCL kernel,
 * src/nanopyx/core/transform/_le_esrrf.pyx:106-163) with one graph launch.
 */
int npx_capture_begin(void **capture_stream);
int npx_capture_end(void *capture_stream, void **graph);
int npx_graph_launch(void *graph, void *stream);
int npx_graph_destroy(void *graph);

/*
 * Serving session (continuous batching of eSRRF + temporal-correlation jobs). No reference counterpart: NanoPyx has
 * no serving layer; every job computes what one timepoint of the eSRRF workflow computes,
 * src/nanopyx/methods/esrrf/eSRRF_workflow.py:98-119 (eSRRF_ST.run, then calculate_eSRRF_temporal_correlations,
 * src/nanopyx/core/transform/sr_temporal_correlations.py:46-75,205-221), with the results npx_esrrf_temporal gives.
 *
 * npx_session_create allocates one device pool of pool_bytes that holds, per admitted job, its geometry tables, its
 * running temporal moments (5 * H*M * W*M floats), partial moments of the frame chunks a slice of a small job runs
 * as, and — for jobs submitted from host memory — its stack (npx_session_job_bytes gives the total). At most
 * max_active jobs are admitted at once; one step aims to process about px_frames_per_step output pixel-frames.
 *
 * npx_session_submit queues a job (never blocks). src is a device pointer (src_is_host 0; it must stay valid until
 * the job is done) or a host pointer, preferably pinned (src_is_host 1; copied into the pool from admission on, in
 * pieces of about 32 MB whose frames become runnable as each piece lands; it must stay valid until the job is done).
 * Outputs are caller-owned device maps (H*M, W*M); those of unset mode bits may be NULL.
 *
 * npx_session_step retires finished jobs, admits queued jobs that fit the pool (arrival order; a job that does not
 * fit yet lets later ones pass), then launches one frame slice of every admitted job whose frames have landed, packed
 * up to 8 jobs per launch, on `stream`. It never blocks; it launches nothing while two earlier steps are still
 * running on the GPU. *launches receives the number of launches made.
 *
 * npx_session_poll returns 0 (queued), 1 (admitted / running) or 2 (done: maps complete); npx_session_forget drops a
 * done job's record. npx_session_stats fills stats[6]: queued jobs, admitted jobs, free pool bytes, launches,
 * slices, steps. npx_session_destroy waits for the device and frees the pool. A session's calls are serialised by
 * its own mutex; one session drives one stream.
 */
int npx_session_create(size_t pool_bytes, int max_active, long long px_frames_per_step, void **session);
int npx_session_submit(void *session, const void *src, int src_is_host, int dtype, int N, int H, int W, int M,
                       double R, double S, int iw, unsigned mode_mask, float *d_avg, float *d_var, float *d_tac2,
                       long long *job_id);
/* A job whose stack the caller copies into device memory itself, piece by piece (d_stack: the device stack, filled
 * over time): npx_session_feed(session, job, frames, stream) announces that frames [0, frames) are on the device
 * once the work queued so far on `stream` (the caller's copies) has completed; frame counts grow with each call and
 * the last one is N. Slices follow the announcements (chunks of at most piece_frames frames). The caller keeps
 * d_stack and the outputs alive until the job is done. */
int npx_session_submit_fed(void *session, const void *d_stack, int dtype, int N, int H, int W, int M, double R,
                           double S, int iw, unsigned mode_mask, float *d_avg, float *d_var, float *d_tac2,
                           int piece_frames, long long *job_id);
int npx_session_feed(void *session, long long job_id, int frames, void *stream);
/* Pool bytes npx_session_submit would need for such a job (0 for invalid parameters). */
size_t npx_session_job_bytes(int N, int H, int W, int M, double R, int dtype, int src_is_host);
/* Host-only check of a job's slicing: lays the job out as npx_session_submit would and slices it as npx_session_step
 * would when it runs alone with px_frames_per_step per step, twice: stepped as each upload piece of a host stack
 * lands (src_is_host 1; all frames at once from device memory), and stepped only after every piece has landed.
 * info[6]: frame chunks a slice may run as (partial slots), frames per chunk, most chunks any slice ran as (never
 * above info[0]), slices of both passes, slices not starting on a chunk boundary (0), upload pieces. No device
 * calls. */
int npx_session_slice_plan(int N, int H, int W, int M, double R, int dtype, int src_is_host, long long px_frames_per_step,
                           int *info);
int npx_session_step(void *session, void *stream, int *launches);
int npx_session_poll(void *session, long long job_id);
int npx_session_forget(void *session, long long job_id);
int npx_session_stats(void *session, long long *stats);
int npx_session_destroy(void *session);

/*
 * SRRF radiality of a device (N, H, W) stack (dtype NPX_DTYPE_F32 or NPX_DTYPE_U16) into d_out (N, H*M, W*M) float32: 12 ring
 * samples of radius ringRadius*M around each output pixel of [(1+border)M, (H-1-border)M) x [(1+border)M,
 * (W-1-border)M), every other output pixel 0. With iw the map is multiplied by d_interp (N, H*M, W*M), or, when
 * d_interp is NULL, by the xM Catmull-Rom magnification of the input computed in the same kernel.
 * Replaces Radiality._run_threaded / _run_opencl, src/nanopyx/core/transform/_le_radiality.pyx:86-126 (per-pixel
 * kernel src/include/_c_sr_radiality.c:15-51, gradients src/include/_c_gradients.c:3-25); with d_interp NULL it is
 * also the CRShiftAndMagnify + Radiality pair of SRRF(), src/nanopyx/methods/srrf/SRRF_workflow.py:50-67.
 */
int npx_radiality(const void *d_img, int dtype, const float *d_interp, int N, int H, int W, int M, double ringRadius,
                  int border, int positivity, int iw, float *d_out, void *stream);

/*
 * Host-only introspection of a radiality plan: info[12] = HM, WM, j_lo, j_hi, i_lo, i_hi, gradient tile rows and
 * columns, output rows and columns per block, row bands, column tiles; ring[25] = ringRadius*M, xRing[12],
 * yRing[12]. No reference counterpart: it exists so the host can check the ring and the launch geometry.
 */
int npx_radiality_info(int H, int W, int M, double ringRadius, int border, int *info, float *ring);

/*
 * Host-only copy of a radiality plan's tables: row (H*M x 80 floats), col (W*M x 80), tile_r0 (row bands),
 * tile_c0 (column tiles); per entry the 12 ring samples' first tap (int bits, tile-relative), 4 weights each, the
 * offsets xc - x0 (or yc - y0), then the intensity taps. No reference counterpart: it lets the host check the
 * tables against the reference's arithmetic without a GPU.
 */
int npx_radiality_tables(int H, int W, int M, double ringRadius, int border, float *row, float *col, int *tile_r0,
                         int *tile_c0);

/*
 * eSRRF 3D of a device (F, S, R, C) stack (dtype NPX_DTYPE_F32 or NPX_DTYPE_U16) into d_out (S*Mz, R*Mxy, C*Mxy)
 * float32: per frame the xMxy Catmull-Rom of the intensity and of the 3D gradients (z slices repeated Mz times), the
 * radial gradient convergence of every voxel of [int(radius_z*Mz), S*Mz - int(radius_z*Mz)) x [int(radius*Mxy), ...)^2
 * with the sensitivity power and, with iw, the intensity weight; frames folded into their running mean (mode 0,
 * "average") or standard deviation (mode 1, "std"). d_work is npx_esrrf3d_workspace() bytes of device scratch.
 * Replaces eSRRF3D._run_threaded / _run_opencl, src/nanopyx/core/transform/_le_esrrf3d.pyx:73-175 (per-voxel kernel
 * src/include/_c_sr_radial_gradient_convergence.c:132-204, gradients src/include/_c_gradients.c:75-116,
 * interpolate_3d_zlinear src/nanopyx/core/transform/_interpolation.pyx:54-103).
 */
size_t npx_esrrf3d_workspace(int F, int S, int R, int C, int Mxy, int Mz, int mode);
/* Workspace bytes per chunk of frames that later npx_esrrf3d calls size their chunks by (0 leaves it); returns the
 * previous value. No reference counterpart: the chunking is this library's. */
size_t npx_esrrf3d_chunk_bytes(size_t bytes);
int npx_esrrf3d(const void *d_img, int dtype, int F, int S, int R, int C, int Mxy, int Mz, double radius,
                double radius_z, double sensitivity, int mode, int iw, float *d_out, void *d_work, size_t work_bytes,
                void *stream);

/*
 * Roberts cross gradients of a device (N, H, W) stack (f32 or u16) into d_gc, d_gr (N, H, W) float32, the previous
 * row and column clamped at 0. Replaces GradientRobertsCross._run_threaded,
 * src/nanopyx/core/transform/_le_roberts_cross_gradients.pyx:67-83 (per pixel src/include/_c_gradients.c:53-72).
 */
int npx_roberts(const void *d_img, int dtype, int N, int H, int W, float *d_gc, float *d_gr, void *stream);

/*
 * Fourier interpolation of a device (N, h, w) float32 stack by an integer factor M into d_out (N, h*M, w*M): frames
 * larger than tile_size (0 = 256) on either side are split into tiles with `overlap` pixels on each side, each
 * interpolated with mirror padding and blended with a Hann window; smaller frames are interpolated whole, mirror
 * padded when `mirror`. d_work is npx_fht_workspace() bytes of device scratch. Replaces fht_space_interpolation /
 * fht_space_interpolation_tiled, src/nanopyx/core/transform/interpolation.py:58-292 (the use_fht branch of eSRRF,
 * src/nanopyx/core/transform/_le_esrrf.pyx:183-188).
 */
size_t npx_fht_workspace(int N, int h, int w, int M, int mirror, int tile_size, int overlap);
int npx_fht_interpolate(const float *d_in, int N, int h, int w, int M, int mirror, int tile_size, int overlap,
                        float *d_out, void *d_work, size_t work_bytes, void *stream);
/*
 * npx_fht_interpolate at the lattice positions only, compact: d_out (N, LH, LW) holds the interpolated (N, h*M, w*M)
 * result at output rows that are multiples of `stride` (>= 2) and the last row, and likewise for columns, in order;
 * LH = ceil(h*M / stride) + (1 if (h*M - 1) % stride != 0 else 0), LW likewise. The rescale still uses the min / max
 * of the whole interpolated grid, so every value equals npx_fht_interpolate's at that position. For eSRRF's use_fht
 * branch (_le_esrrf.pyx:183-188): its radial gradient convergence reads the x2M gradients only at positions k*M
 * clamped to the last row / column (_c_sr_radial_gradient_convergence.c:31-32,45-61); stride = M gives the
 * (2H + 1, 2W + 1) lattice npx_rgc_lattice reads. Same workspace as npx_fht_interpolate.
 */
int npx_fht_interpolate_lattice(const float *d_in, int N, int h, int w, int M, int mirror, int tile_size, int overlap,
                                int stride, float *d_out, void *d_work, size_t work_bytes, void *stream);

/*
 * Catmull-Rom shift and magnify of a device (N, H, W) stack (f32 or u16) into d_out (N, int(H*mag_row),
 * int(W*mag_col)) float32: output (i, j) samples the input at (i / mag_row - shift_row, j / mag_col - shift_col),
 * 0 outside it. Replaces ShiftAndMagnify._run_threaded (CRShiftAndMagnify),
 * src/nanopyx/core/transform/_le_interpolation_catmull_rom.pyx:143-170 (per pixel
 * src/include/_c_interpolation_catmull_rom.c:21-52). npx_shift_magnify_shape gives the output size.
 */
int npx_shift_magnify_shape(int H, int W, float mag_row, float mag_col, int *HM, int *WM);
int npx_shift_magnify(const void *d_img, int dtype, int N, int H, int W, float shift_row, float shift_col,
                      float mag_row, float mag_col, float *d_out, void *stream);

/*
 * In-place macro-pixel correction of a device (S, HM, WM) stack; d_cmap is S*M*M floats of scratch.
 * Replaces macro_pixel_corrector, src/nanopyx/core/transform/mpcorrector.pyx:7-73.
 */
int npx_macro_pixel(float *d_img, int S, int HM, int WM, int M, float *d_cmap, void *stream);

/*
 * Host-only introspection of the geometry tables the kernels use (no GPU needed). info[16] receives
 * K, margin, HM, WM, nphase, Q (rows per lane) of the per-frame kernel for this shape, Q of the temporal kernel, floats per
 * phase ring block, HG, HI (tile halos), LP, waves per block, phase groups, prefetch slots, 0, 0; finfo[4] receives
 * fwhm, sigma, tSS, tSO. No reference counterpart: it exists so the host can check launch geometry and tables.
 */
int npx_plan_info(int H, int W, int M, double R, int *info, float *finfo);

/* Copies the plan tables to host arrays sized from npx_plan_info (ring nphase*ring_block floats,
 * wsum nphase, lattice taps 2H+1 / 2W+1 entries, intensity taps HM / WM entries; *_w arrays hold 4 floats
 * per entry). The ring blocks and wsum are the exact per-phase geometry; the kernels form each block's ring
 * geometry from its float32 pixel centres, equal to these at power-of-two magnifications. No reference
 * counterpart (introspection, as npx_plan_info). */
int npx_plan_tables(int H, int W, int M, double R, float *ring, float *wsum, int *lrb, float *lrw, int *lcb,
                    float *lcw, int *crb, float *crw, int *ccb, float *ccw);

/* Launch shape of the fused kernels for subsequent calls: at most max_waves_per_block output phases (one
 * wavefront each) per block, and the block count the temporal kernel aims for when it splits the frames into
 * chunks (0 = defaults 16 / 2048). No reference counterpart: it is what the autotuner varies. Changing it
 * between calls changes the frame chunking and therefore the last bits of the temporal moments. */
int npx_set_launch(int max_waves_per_block, int target_blocks);

/*
 * Autotuner of the launch shape, per (H, W, M, R). npx_autotune times every candidate shape — waves (output phases)
 * per block from {16, 8, 4} and a target block count from {256 .. 8192}, which sets the frames per chunk — with
 * npx_esrrf_temporal on the caller's stack and outputs (HIP events on `stream`; blocking; after >= 20 ms of warm-up
 * calls, each candidate's best of three interleaved batches of >= ~2 ms), keeps the fastest for later calls with the
 * same (H, W, M, R) and returns its time per call in *best_ms. Candidates whose chunking needs
 * more than work_bytes are skipped; npx_autotune_workspace gives the workspace that admits all of them.
 * npx_tune_set stores a shape directly (0, 0 forgets it); npx_tune_get writes the shape calls with these parameters
 * use into cfg[2] and returns 1 if it was tuned, 0 for the defaults. No reference counterpart: it takes the place
 * of the OpenCL path's fixed frames-per-chunk sizing from device memory (max_slices,
 * src/nanopyx/core/transform/_le_esrrf.pyx:64-69) and of the Liquid Engine's run-type benchmarking.
 */
size_t npx_autotune_workspace(int N, int H, int W, int M, double R);
int npx_autotune(const void *d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                 unsigned mode_mask, float *d_avg, float *d_var, float *d_tac2, void *d_work, size_t work_bytes,
                 void *stream, double *best_ms);
int npx_tune_set(int H, int W, int M, double R, int max_waves_per_block, int target_blocks);
int npx_tune_get(int H, int W, int M, double R, int *cfg);
/*
 * The same for the per-frame-map kernels (npx_esrrf_frames, and npx_rgc / npx_rgc_lattice on an H x W image magnified
 * M times), whose best shape can differ from the temporal kernel's: npx_autotune_frames times every candidate with
 * npx_esrrf_frames on the caller's stack into d_out ((N, H*M, W*M) float32; blocking) and keeps the fastest for
 * (H, W, M, R); npx_tune_frames_set / npx_tune_frames_get store and read such a shape (get returns 1 when the
 * per-frame kernels have their own, 0 when they use the temporal kernel's shape, which they do until tuned). The
 * maps do not depend on the shape. No reference counterpart (as npx_autotune).
 */
int npx_autotune_frames(const void *d_in, int dtype, int N, int H, int W, int M, double R, double S, int iw,
                        float *d_out, void *stream, double *best_ms);
int npx_tune_frames_set(int H, int W, int M, double R, int max_waves_per_block, int target_blocks);
int npx_tune_frames_get(int H, int W, int M, double R, int *cfg);

/* Turns per-call HIP-event timing of the eSRRF kernels on (1) or off (0). No reference counterpart: it lets a
 * caller (the benchmark) read the fused kernel's own duration on the stream it was launched on. */
int npx_timing(int on);

/* Waits for and sums the events recorded since the last read: number of calls, total milliseconds of the fused
 * eSRRF kernel and of the moment merge that follows it. No reference counterpart (see npx_timing). */
int npx_timing_read(int *n_calls, double *kernel_ms, double *merge_ms);

/* Frees the cached device tables and FFT plans of every device (optional; the process exit frees them too). No call
 * may be in flight. */
void npx_release(void);

#ifdef __cplusplus
}
#endif

#endif /* NPX_ESRRF_H */
