/*
 * ose_api.h -- the extern "C" boundary of the MI355X OSE time stepper.
 *
 * The reference (Littleke/3D-OSE) has no plugin or operator API: its hot
 * path is the C function wavefield_extro_ori2(cufftHandle, cufftResult)
 * (ose_3d_lib.cu:639) reading ~20 program globals (ose_3d.cu:42-62), fed by
 * Xt_add_wavelet_sxyz (ose_3d_lib.cu:610) and read by rec_shotdata_3d
 * (ose_3d_kernel.cu:412) + rec_slice_to_data (ose_3d_lib.cu:876).  This
 * header is what that path's callers bind instead.  Each entry point names
 * the reference code it replaces.
 *
 * Conventions
 *   - Arrays are [z][y][x] with x fastest (ose_3d_kernel.cu:35); "padded"
 *     extents include the absorbing pads.
 *   - Host pointers are caller-owned and only read or written during the
 *     call; the library owns all device memory.
 *   - Every function returns an OSE_* status; nothing exits the process.
 *     ose_last_error() gives the text of the last failure on a context.
 *   - One host thread per context; ose_run is asynchronous on the
 *     context's HIP stream, getters synchronize.
 */
#ifndef OSE_API_H
#define OSE_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSE_OK            0
#define OSE_EINVAL        1   /* bad argument                                  */
#define OSE_ENOMEM        2   /* device or host allocation failed              */
#define OSE_EDEVICE       3   /* HIP runtime error                             */
#define OSE_ESTATE        4   /* call not valid in the context's current state */
#define OSE_ENUMERIC      5   /* non-finite input                              */
#define OSE_EUNSUPPORTED  6   /* FFT extent outside [2, 4096]                  */

typedef struct ose_ctx ose_ctx;

/* One velocity model on one padded grid.  Replaces the globals
 * nxb,nyb,nzb, dkx..dkz, R, dtR, M and att_blr/att_bfb/att_bud that
 * wavefield_extro_ori2 reads (ose_3d.cu:42-62, 120-122, 191-199;
 * Alloc_mulGPU ose_3d_lib.cu:123-159).                                     */
typedef struct {
    int nxb, nyb, nzb;     /* padded extents = FFT sizes                       */
    int bl, bf, bu;        /* interior origin inside the padded grid           */
    int nx, ny, nz;        /* interior extents                                 */
    double dx, dy, dz, dt; /* grid spacing (m), time step (s)                  */
    double R;              /* vmax*pi*sqrt(dx^-2+dy^-2+dz^-2)  (ose_3d.cu:194) */
    int K;                 /* highest Bessel order = operator applications per
                              step, >= 2 (reference: M+1, ose_3d.cu:197)       */
    const double* coef;    /* K+1 values: J0(R dt), 2 J_k(R dt), k = 1..K      */
    const float* taper_x;  /* nxb / nyb / nzb accumulator taper profiles       */
    const float* taper_y;  /*   (generate_att, ose_3d_kernel.cu:242-261)       */
    const float* taper_z;
    int rec_z;             /* receiver plane: interior depth index (ref: 0)    */
} ose_model_desc;

/* Context on one HIP device.  Replaces Alloc/Alloc_cufftXt/cufftXtSetGPUs/
 * cufftMakePlan3d (ose_3d_lib.cu:47-105,162-184; ose_3d.cu:207-237).     */
int ose_create(ose_ctx** ctx, int device);
/* Replaces Free_cufftXt/cufftDestroy/Free (ose_3d_lib.cu:186-250).         */
int ose_destroy(ose_ctx* ctx);
const char* ose_last_error(const ose_ctx* ctx);

/* Register a model.  v_zyx: padded velocity [nzb][nyb][nxb] in m/s (the
 * caller edge-pads as add_pml_layers_v_h, ose_3d_lib.cu:912-1032, does).
 * Replaces the H2D of vbD (ose_3d.cu:236) and the taper set-up.            */
int ose_add_model(ose_ctx* ctx, const ose_model_desc* desc, const float* v_zyx,
                  int* model_id);

/* Register a shot on a model: source at padded cell (sx, sy, sz), wavelets
 * s (added to Re u) and h (added to Im u) of length nt, injected before
 * each step (Xt_add_wavelet_sxyz ose_3d_lib.cu:610-633, add_wavelet_Xt
 * ose_3d_kernel.cu:291-297).  Traces are kept for nt steps.                */
int ose_add_shot(ose_ctx* ctx, int model_id, int sx, int sy, int sz,
                 const float* s, const float* h, int nt, int* shot_id);

/* Enqueue nsteps OSE time steps for every shot of every model (the body of
 * the time loop ose_3d.cu:248-275: inject, wavefield_extro_ori2,
 * rec_shotdata_3d).  Asynchronous; the first call allocates the state.   */
int ose_run(ose_ctx* ctx, int nsteps);
int ose_sync(ose_ctx* ctx);
/* Same as ose_run + ose_sync, timed with HIP events on the context stream. */
int ose_timed_run(ose_ctx* ctx, int nsteps, double* ms);

/* Receivers of a shot at listed padded cells xyz[3r..3r+2] = (x, y, z),
 * r < n, instead of the plane rec_z.  Replaces the receiver geometry the
 * reference builds but does not use: gxbeg..gzbeg, jgx..jgz, ng
 * (ose_3d.cu:100-114) through sg_init_device (ose_3d_kernel.cu:202-212)
 * and the commented-out cell lookup of rec_shotdata_3d
 * (ose_3d_kernel.cu:417-423).  Before the first ose_run; n = 0 restores
 * the plane.  Receivers outside the grid fail with OSE_EINVAL.            */
int ose_set_receivers(ose_ctx* ctx, int shot_id, const int* xyz, int n);
/* Spatially extended source for a shot, before the first ose_run: each
 * step adds (s[it] + i h[it]) px[x-x0] py[y-y0] pz[z-z0] over the padded box
 * box = {x0, y0, z0, wx, wy, wz} (widths 1..64, inside the grid) instead of
 * the point of ose_add_shot (Xt_add_wavelet_sxyz ose_3d_lib.cu:610-633,
 * add_wavelet_Xt ose_3d_kernel.cu:291-297).  A point injected every dt
 * excites grid modes above the temporal Nyquist pi/dt, whose aliases spoil
 * large time steps; a profile band-limited below it does not.  box = NULL
 * restores the point.  No counterpart in the reference.                   */
int ose_set_source_profile(ose_ctx* ctx, int shot_id, const int* box, const float* px,
                           const float* py, const float* pz);
/* Number of traces of a shot: its listed receivers, else nx*ny.           */
int ose_receiver_count(ose_ctx* ctx, int shot_id, int* nrec);

/* Receiver traces Im(u) on the plane rec_z over the interior x, y (or at
 * the shot's listed receivers, in list order), for steps [it0, it0+nsteps):
 * out is trace-major [iy*nx+ix][nsteps] like the reference's seismogram
 * (rec_shotdata_3d ose_3d_kernel.cu:412-436, rec_slice_to_data
 * ose_3d_lib.cu:876-884, write ose_3d.cu:276).                            */
int ose_get_traces(ose_ctx* ctx, int shot_id, int it0, int nsteps, float* out);
/* Upload recorded traces (same layout as ose_get_traces), e.g. when a run
 * resumes from a checkpoint.  No counterpart in the reference.            */
int ose_set_traces(ose_ctx* ctx, int shot_id, int it0, int nsteps, const float* in);

/* ---- snapshot streaming ------------------------------------------------
 * The reference dumps the whole padded state after every step, one
 * synchronous copy and fwrite per lane (the commented block ose_3d.cu:258-269,
 * write_3dfile_from_1darray_ci ose_3d_lib.cu:364-380).  Here, every `every`
 * steps, each shot's state is encoded on the device, copied to pinned host
 * memory on a second stream while stepping continues, and passed to `fn` on
 * a library thread.  The box is the padded grid (interior = 0) or the
 * interior; lanes: bit 0 Re u (P), bit 1 Im u (Q).  Encodings:
 *   OSE_SNAP_F32   float32, exact
 *   OSE_SNAP_BF16  bfloat16, round to nearest even
 *   OSE_SNAP_Q16   per x-line of the box a float scale s = max|v| and int16
 *                  q = rint(v * 32767 / s), decoded as q * (s / 32767)
 *   OSE_SNAP_Q8    the same with int8 and 127
 * Data layout: Q formats first hold the scales, float [lane][z][y]; then
 * the values [lane][z][y][x], x fastest, lanes in the order P, Q.  `data`
 * is valid during the call only; fn must not call into the context.
 * ose_sync and ose_snapshot_drain return once every snapshot enqueued so
 * far has been delivered.  every = 0 turns streaming off.  Snapshots are
 * taken after ose_run's steps whose step count (after the step) is a
 * multiple of every.                                                      */
#define OSE_SNAP_F32 0
#define OSE_SNAP_BF16 1
#define OSE_SNAP_Q16 2
#define OSE_SNAP_Q8 3
typedef struct {
    int shot_id;
    int step;                /* the state after `step` steps                 */
    int lanes, format;
    int x0, y0, z0;          /* box origin in the padded grid                */
    int nx, ny, nz;          /* box extents                                  */
    size_t bytes;
} ose_snapshot_info;
typedef void (*ose_snapshot_fn)(void* user, const ose_snapshot_info* info, const void* data);
int ose_set_snapshots(ose_ctx* ctx, int every, int lanes, int format, int interior,
                      ose_snapshot_fn fn, void* user);
int ose_snapshot_drain(ose_ctx* ctx);

/* Current state u = P + iQ of a shot over the padded grid (the p2D/q2D .y
 * lanes; snapshot dump ose_3d.cu:260-270).                                 */
int ose_get_wavefield(ose_ctx* ctx, int shot_id, float* re, float* im);
/* Overwrite the state of a shot (checkpoint restore; initial conditions). */
int ose_set_wavefield(ose_ctx* ctx, int shot_id, const float* re, const float* im);
/* Step counter: index into the wavelets and traces of the next step.       */
int ose_get_step(ose_ctx* ctx, int* it);
int ose_set_step(ose_ctx* ctx, int it);

/* out = (v/R) IFFT3(|k| FFT3(in)) / N for one complex field (interleaved
 * re,im, padded grid of the model): the operator application of
 * ose_3d_lib.cu:649-687 (memcpy_between_device_cr, cufftXtExec FORWARD,
 * scale_kxkz_3D, INVERSE, scal_vr_3d).                                      */
int ose_apply_operator(ose_ctx* ctx, int model_id, const float* in, float* out);

/* ose_apply_operator without any device allocation: the two work fields
 * live in a static device array of the library (64 MiB per GPU, held until
 * the process ends), so models of at most 4194304 padded cells (e.g. 160^3)
 * are accepted and larger ones fail with OSE_EINVAL.  Same inputs, outputs
 * and results as ose_apply_operator, bit for bit.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (several host
 * threads, several contexts) race; serialise them.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc). */
int ose_apply_operator_noalloc(ose_ctx* ctx, int model_id, const float* in, float* out);

/* Unnormalised 3D DFT of an interleaved complex [nz][ny][nx] array with the
 * library's pass kernels; sign -1 forward, +1 inverse (cufftXtExecDescriptorC2C,
 * ose_3d_lib.cu:668).                                                       */
int ose_fft3d(ose_ctx* ctx, int nx, int ny, int nz, const float* in, float* out, int sign);

/* use_graph: replay one captured hipGraph per step (default 1).
 * profile:   record HIP events around every pass kernel (eager launches)
 *            so ose_pass_stats can report per-pass durations.              */
int ose_set_options(ose_ctx* ctx, int use_graph, int profile);
/* Streams a step runs on: 2 (default) deals the shots of a context with two
 * or more models to two lanes that run side by side (a model of two or more
 * shots in halves, the second lane starting after the first lane's
 * injection, forward-y and z pass); 1 runs
 * every launch in order on one stream.  A single model's shots always run
 * in one lane, and per-pass profiling (ose_set_options profile = 1) always
 * uses one stream.  No counterpart in the reference (its shots run one
 * after another, ose_3d.cu:242).                                          */
int ose_set_streams(ose_ctx* ctx, int n);
/* Per-pass statistics since the last reset: pass 0 forward-y, 1 z-symbol,
 * 2 inverse-y, 3 x-epilogue (middle orders that update the accumulator),
 * 4 x-epilogue (first order), 5 x-epilogue (last order), 6 injection,
 * 7 x-epilogue (middle orders that leave the accumulator alone; the step
 * updates it on about half of the orders, adding two terms each time),
 * 8 x-epilogue (the middle order that starts the accumulator: no A read).
 * avg_ms is the mean HIP-event
 * duration of one launch, bytes the mean algorithmic HBM bytes of one
 * launch (launches of several models differ in size), so bytes / avg_ms
 * is the pass's aggregate rate.                                             */
int ose_pass_stats(ose_ctx* ctx, int pass, double* avg_ms, long* launches, double* bytes);
int ose_reset_stats(ose_ctx* ctx);

/* Launch-shape autotuner (default on): at the first ose_run each pass of each
 * model is timed in every compiled shape (x columns per workgroup for the
 * strided passes, lines per workgroup for the x pass) and the fastest kept.
 * No counterpart in the reference (cuFFT chose its own plan, ose_3d.cu:226). */
int ose_set_autotune(ose_ctx* ctx, int on);
/* shape[4] = {T forward-y, T z-symbol, T inverse-y, x-pass code} of a
 * model; x-pass code = P + 1000 * b where b = 1 marks the variant with a
 * scheduling barrier before the twiddle multiplies; a negative value marks
 * the half-LDS (split real/imaginary) exchange.  The x code is the shape of
 * the middle orders that update the accumulator; the autotuner times the
 * step's other x-pass kinds (orders that leave the accumulator alone, its
 * first update, the first and the last order) separately and may run each
 * in another shape.                                                       */
int ose_get_tuning(ose_ctx* ctx, int model_id, int* shape);
/* Pin a model's launch shapes (same encoding as ose_get_tuning; the
 * strided entries must be compiled for the y and z extents; the x code
 * applies to every x pass).  Skips the autotuner for that model.  No
 * counterpart in the reference.                                           */
int ose_set_tuning(ose_ctx* ctx, int model_id, const int* shape);
/* Every launch shape of a model, one entry per pass kind: shape[8] =
 * {T forward-y, T z-symbol, T inverse-y, x code of the middle orders that
 * update the accumulator, of those that leave it alone, of its first
 * update, of the first order, of the last order} (codes as ose_get_tuning).
 * ose_set_tuning_modes pins all eight (skipping the autotuner for that
 * model), so a configuration read back from one context reproduces its
 * launches exactly in another.  No counterpart in the reference.         */
int ose_get_tuning_modes(ose_ctx* ctx, int model_id, int* shape);
int ose_set_tuning_modes(ose_ctx* ctx, int model_id, const int* shape);
/* The compiled launch shapes of a model: strided codes usable on both the y
 * and z passes, and x-pass codes.  On entry *nstrided, *nxs are the array
 * capacities; on return, the counts.  No counterpart in the reference.   */
int ose_launch_shapes(ose_ctx* ctx, int model_id, int* strided, int* nstrided, int* xshapes, int* nxs);

/* Bandwidth yardstick: a plain 16-byte-per-lane device copy of `bytes`,
 * repeated reps times; gbs = (read + write bytes) / time.  No counterpart in
 * the reference; it calibrates the per-pass GB/s and the HBM counters.    */
int ose_copy_probe(ose_ctx* ctx, size_t bytes, int reps, double* gbs);
/* Access-pattern bound of a strided pass: the HBM reads and writes of model
 * model_id's forward-y (axis 1) or z-symbol (axis 2) pass over all its
 * shots, from one of the model's two spectral work fields to the other
 * (the state is intact between steps), in the pass's current launch shape
 * (same tile order, addressing and workgroups per CU; for a chunked z pass
 * its one-tile counterpart of the same width) but without the
 * transforms, repeated reps times; gbs = (read + write bytes) / time.  The
 * pass's own rate over this is the share of its access pattern's bound it
 * reaches.  Compiled extents only (OSE_EUNSUPPORTED on the runtime-length
 * path).  No counterpart in the reference.                               */
int ose_pattern_probe(ose_ctx* ctx, int model_id, int axis, int reps, double* gbs);

/* Mean duration (ms) of one launch of a single pass over all shots of a
 * model, on scratch buffers: which = 0 forward-y, 1 plain forward-z,
 * 2 z-symbol, 3 inverse-y, 4 x-epilogue (middle order), 5 forward-x.
 * Measurement utility; no counterpart in the reference.                   */
int ose_pass_probe(ose_ctx* ctx, int model_id, int which, int reps, double* ms);

/* ---- learned thin pad ------------------------------------------------
 * A 3-layer 3D convolution (9 input features -> 16 -> 16 -> 2 channels,
 * 3x3x3 taps, GELU; BF16 weights, FP32 accumulation on MFMA) that corrects
 * the pad cells of the new state after every step with a learned passive
 * complex gain: u' = u G, G = q / max(1, |q|), q = 1 + y, y = net(u/s, v,
 * distance into the pad), s the RMS of u on the network's receptive shell.
 * It replaces the reference's wide absorbing layer (generate_att /
 * apply_ABC_Xt, ose_3d_kernel.cu:242-288) with a thin one; the classical
 * taper of the model stays active underneath.  Parameters: ose_boundary_net_size()
 * floats, [W1[27][16][16] (tap, out, in), b1[16], W2, b2, W3, b3].          */
int ose_boundary_net_size(void);
/* Enable the network on a model (params == NULL removes it).  guard > 1:
 * if one step would multiply the pad energy by more than guard (or produce
 * a non-finite value), the correction is not applied and the model falls
 * back to the plain taper for the rest of the run.  replace_taper != 0: while
 * the network is on, the model's taper is 1 (the network is the absorbing
 * layer); the fallback restores the classical taper.  Non-finite parameters
 * are rejected with OSE_ENUMERIC (the taper stays).                        */
int ose_set_boundary_net(ose_ctx* ctx, int model_id, const float* params, size_t n, double guard,
                         int replace_taper);
/* active: the network is in use; tripped: its guard has fired.             */
int ose_boundary_status(ose_ctx* ctx, int model_id, int* active, int* tripped);
/* Put a model's network into the tripped state from the host, as if its
 * guard had fired: the plain taper from the next step on, status tripped.
 * A run resumed from a checkpoint taken after a trip restores it this way,
 * so it continues as the uninterrupted run would.  No counterpart in the
 * reference.                                                              */
int ose_trip_boundary_net(ose_ctx* ctx, int model_id);

/* Training object of the network on one padded grid (single sample at a
 * time; BF16 forward and backward on MFMA, FP32 Adam).  Fields u, t are
 * interleaved complex [nzb][nyb][nxb]; v_padded in m/s.                   */
typedef struct ose_bnet ose_bnet;
int ose_bnet_create(ose_bnet** net, int device, int nxb, int nyb, int nzb, int bl, int bf, int bu,
                    int nx, int ny, int nz, const float* v_padded);
int ose_bnet_destroy(ose_bnet* net);
const char* ose_bnet_error(const ose_bnet* net);
int ose_bnet_set_params(ose_bnet* net, const float* params, size_t n);
int ose_bnet_get_params(ose_bnet* net, float* params, size_t n);
/* delta = u (G - 1) on the pad cells, zero elsewhere (interleaved).      */
int ose_bnet_forward(ose_bnet* net, const float* u, float* delta);
/* loss = mean over pad cells of |(u/s) G - t/s|^2 and its gradient.     */
int ose_bnet_gradient(ose_bnet* net, const float* u, const float* t, float* loss, float* grad);
/* Keep a training pair on the device.                                     */
int ose_bnet_add_sample(ose_bnet* net, const float* u, const float* t, int* count);
/* iters Adam steps on randomly drawn stored pairs; losses[iters].          */
int ose_bnet_train(ose_bnet* net, int iters, float lr, unsigned long long seed, float* losses);

/* FFT extents.  The reference plans its transform for any padded extent
 * (cufftMakePlan3d(nzb, nyb, nxb), ose_3d.cu:226).  Here every extent in
 * [2, 4096] runs: the extents N = N1 N2 whose factors both fit the
 * compiled register DFTs take those kernels, every other one (large prime
 * factors, extents above 1024, ...) a runtime-length mixed-radix path with
 * the same passes and results (slower: no register-resident transform).
 * ose_extent_kind: 1 compiled in several launch shapes (the autotuner picks
 * one), 3 compiled in a single launch shape, 2 the runtime-length path, 0
 * not supported.  ose_supported_extent: the smallest extent >= n of kind 1,
 * or 0 (what the host's extent=smooth rounds up to).                      */
int ose_supported_extent(int n);
int ose_extent_is_supported(int n);
int ose_extent_kind(int n);
/* path 1: every model added afterwards (and ose_fft3d) uses the
 * runtime-length kernels on all three axes, compiled or not; 0 (default):
 * compiled kernels where they exist.  Before the first ose_add_model.  For
 * tests and for comparing the two paths; no counterpart in the reference. */
int ose_set_fft_path(ose_ctx* ctx, int path);
/* The step's accumulator schedule for K orders (host only, no device).  The
 * series sum A = sum_m c_m g^(e_m) W_m of wavefield_extro_ori2
 * (ose_3d_lib.cu:741-859: add_bessel_3d then apply_ABC_Xt on every order
 * j >= 1) is formed by only some orders' x passes, each adding the terms it
 * holds (W_{j-1} and W_{j+1}).  plan[6*j ..] = {touch, adds W_{j-1} (W_0 at
 * j = 0), adds W_{j+1} (W_1 at j = 0), ea, eb, ec}: the pass sets
 * A <- g^ea A + g^eb c_{j-1} W_{j-1} + g^ec c_{j+1} W_{j+1} (ea < 0: no
 * earlier A).  K in [2, 64].                                               */
int ose_accumulation_plan(int K, int* plan);

#ifdef __cplusplus
}
#endif
#endif /* OSE_API_H */
