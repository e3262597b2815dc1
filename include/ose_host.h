/*
 * ose_host.h -- C++ host library of the OSE driver (libose_host.so).
 *
 * Everything the reference's main() and ose_3d_lib.cu do on the host (or in
 * <<<1,1>>> kernels) around the time step, in FP64: grid and pad rules,
 * Ricker + Hilbert source, Bessel coefficients, order rule, taper profiles,
 * velocity file I/O and padding, seeded models, seismogram/.su/snapshot
 * output, and a simulation object that drives include/ose_api.h.  It
 * reaches the GPU only through that C-ABI.
 */
#ifndef OSE_HOST_H
#define OSE_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "ose_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grid ------------------------------------------------------------ */
/* judge_odd (ose_3d_lib.cu:36-44): odd padded extent -> after-pad + 1.     */
int oseh_judge_odd(int* after, int* nb);

/* ---- source ---------------------------------------------------------- */
/* generate_wavelet_gpu (ose_3d_kernel.cu:3-13), FP64.                     */
void oseh_ricker(double* w, int nt, double fm, double dt, double amp, int delay);
/* R2C -> hilbert_1d -> C2R (ose_3d_lib.cu:575-583, ose_3d_kernel.cu:154-172).
 * normalize = 0: the reference's per-shot copy (times nt, quirk Q4);
 * normalize = 1: main()'s scaled copy (scale_wavelet, ose_3d.cu:179).      */
void oseh_hilbert(const double* s, double* h, int nt, int normalize);

/* ---- Bessel / order -------------------------------------------------- */
/* J_n(x), n = 0..nmax, by Miller's backward recursion normalised with
 * J0 + 2 sum J_2k = 1 (replaces jnf, ose_3d_kernel.cu:310,388).            */
int oseh_bessel_all(double x, int nmax, double* J);
/* coef[0] = J0(z), coef[k] = 2 J_k(z), k = 1..K.                           */
int oseh_coefficients(double z, int K, double* coef);
/* R = vmax * pi * sqrt(dx^-2 + dy^-2 + dz^-2)  (ose_3d.cu:194).            */
double oseh_spectral_radius(double vmax, double dx, double dy, double dz);
/* Reference order rule: M = (int)(R dt) + 6, K = M + 1 (ose_3d.cu:197).
 * -1 for z = R dt outside [0, 1e4] (or not a number).                      */
int oseh_order_reference(double z);
/* Smallest K with 2 sum_{k>K} |J_k(z)| < tol; -1 for z outside [0, 1e4]
 * or tol <= 0.                                                              */
int oseh_order_tail(double z, double tol);

/* ---- taper ----------------------------------------------------------- */
#define OSEH_TAPER_REF    0   /* generate_att (ose_3d_kernel.cu:242-261)      */
#define OSEH_TAPER_CERJAN 1   /* exp(-(alpha d)^2), d = 1..b from interior    */
#define OSEH_TAPER_NONE   2
void oseh_taper(float* a, int before, int nn, int after, int mode, double alpha);

/* ---- velocity -------------------------------------------------------- */
/* Read a reference velocity file ([x][y][z], z fastest, float32 LE;
 * input_file_xyz_boundary ose_3d_lib.cu:887-909) into interior [z][y][x]. */
int oseh_read_velocity_xyz(const char* path, int nx, int ny, int nz, float* v_zyx);
/* Edge-replicate an interior cube into the padded grid
 * (add_pml_layers_v_h ose_3d_lib.cu:912-1032).                            */
void oseh_pad_velocity(const float* v_int, int nx, int ny, int nz, int bl, int br, int bf,
                       int bb, int bu, int bd, float* v_pad);
/* Seeded generators, interior [nz][ny][nx]:
 *   kind 0: constant v0
 *   kind 1: two layers, v0 above depth index `iface`, v1 below
 *   kind 2: 8 flat jittered layers 1500..4500 m/s + 6 Gaussian anomalies
 *           (sigma 15 cells, +-15 %), 5-cell box smoothing
 *   kind 3: 12 jittered layers 1500..4000 m/s + a dipping ellipsoidal salt
 *           body (4500 m/s, semi-axes 150/120/80 cells scaled to the grid)
 *   kind 4: layered random model for boundary training (3-9 layers,
 *           0-4 anomalies)                                                 */
int oseh_model(int kind, int nx, int ny, int nz, uint64_t seed, double v0, double v1,
               int iface, float* v_zyx);
uint64_t oseh_rng_next(uint64_t* state);   /* xoshiro256** over a 4-word state */
double oseh_rng_uniform(uint64_t* state);
double oseh_rng_normal(uint64_t* state);
void oseh_rng_seed(uint64_t* state, uint64_t seed);

/* ---- output ---------------------------------------------------------- */
/* Raw trace-major float32 seismogram [nrec][nt] (write_3dfile_from_1darray,
 * ose_3d_lib.cu:327-343, called at ose_3d.cu:276).                         */
int oseh_write_raw(const char* path, const float* data, size_t n);
/* SU file: per trace a 240-byte SEG-Y trace header (native endianness:
 * tracl @0, tracr @4, fldr @8, tracf @12, ep @16, cdp @20, offset @36,
 * gelev @40, sdepth @48, scalel @68, scalco @70, sx @72, sy @76, gx @80,
 * gy @84, ns @114 (u16), dt @116 (u16, microseconds)) followed by nt
 * float32 samples.  Coordinates and elevations in metres times 100
 * (scalco = scalel = -100); gelev = -(receiver depth).  rec_x/rec_y/rec_z:
 * receiver coordinates (m, z = depth; any may be NULL for 0), src_x/y/z.  */
int oseh_write_su(const char* path, const float* traces, int nrec, int nt, double dt,
                  const double* rec_x, const double* rec_y, const double* rec_z, double src_x,
                  double src_y, double src_z, int shot_index);
/* Read back an SU file written above: returns nrec, nt and the samples.   */
int oseh_read_su(const char* path, float* traces, size_t cap, int* nrec, int* nt, double* dt);
/* SEG-Y rev 1 file: 3200-byte EBCDIC textual header, 400-byte big-endian
 * binary header (sample interval @3216, samples per trace @3220, format 5 =
 * IEEE float @3224, revision 0x0100 @3500), then per trace the same
 * 240-byte header as the SU file in big-endian order and nt big-endian
 * IEEE float32 samples.  nt and the interval in microseconds must fit in
 * 15 bits.                                                                 */
int oseh_write_segy(const char* path, const float* traces, int nrec, int nt, double dt,
                    const double* rec_x, const double* rec_y, const double* rec_z, double src_x,
                    double src_y, double src_z, int shot_index);
/* Read a SEG-Y file (format 5 IEEE or format 1 IBM floats, fixed trace
 * length): returns nrec, nt, dt and the samples as IEEE float32.          */
int oseh_read_segy(const char* path, float* traces, size_t cap, int* nrec, int* nt, double* dt);
/* Wavefield snapshot over the padded grid, file order [y][x][z] with z
 * fastest (write_3dfile_from_1darray_ci ose_3d_lib.cu:364-380).            */
int oseh_write_snapshot(const char* path, const float* f_zyx, int nxb, int nyb, int nzb);

/* Encoded snapshots (ose_set_snapshots layouts).  Bytes of one lane of an
 * nx x ny x nz box (0 for a bad format or box).                            */
size_t oseh_snapshot_lane_bytes(int format, int nx, int ny, int nz);
/* Copy lane `lane` (0-based among the nl lanes delivered) of a delivered
 * snapshot into out (oseh_snapshot_lane_bytes bytes): its scales, if any,
 * then its values.                                                         */
int oseh_snapshot_lane(int format, int nl, int lane, int nx, int ny, int nz, const void* data, void* out);
/* Decode one lane to float32 [z][y][x].                                    */
int oseh_decode_snapshot(int format, int nx, int ny, int nz, const void* lane_data, float* out);
/* .osnap file: 64-byte header ("OSESNAP1", then int32 format, lane (0 = P,
 * 1 = Q), step, shot, x0, y0, z0, nx, ny, nz at byte 8.., uint64 payload
 * bytes at 48) and one lane's encoded data.  box = {x0, y0, z0, nx, ny, nz}.
 * Reading decodes into out (cap floats, [z][y][x]; out may be NULL to read
 * the header only) and returns the ten header ints in meta.               */
int oseh_write_osnap(const char* path, int format, int lane, int step, int shot, const int* box,
                     const void* lane_data);
int oseh_read_osnap(const char* path, float* out, size_t cap, int* meta);

/* Normalised Gaussian source profile along one axis: exp(-(j-c)^2/(2 s^2))
 * over j in [c - ceil(4s), c + ceil(4s)] clipped to [0, n), summing to 1
 * (s = 0: the unit point).  Writes the first index to *start and the
 * values to out; returns the width, or -1 if it exceeds cap.              */
int oseh_gaussian_profile(double sigma, int center, int n, int* start, float* out, int cap);

/* ---- simulation ------------------------------------------------------ */
/* A simulation is built from SU-style "key=value" parameters (the reference
 * hard-codes them in main(), ose_3d.cu:75-125) or a preset:
 *   preset=ref    the reference as shipped (101^3 from a velocity file or the
 *                 equivalent 3000/4000 two-layer model, 25-cell pads, nt=300)
 *   preset=c0..c4 the five benchmark configurations.
 * Keys: nx ny nz dx dy dz dt nt fm amp delay boundary bl br bf bb bu bd
 *       model seed v0 v1 iface vfile sx sy sz nshot jsx jsy jsz srcgrid
 *       rec_z korder ktol taper alpha extent hilbert device graph
 *       streams (1 or 2, ose_set_streams) max_shots (cap on shots per
 *       model) snap segy src_sigma
 * Receivers: by default the whole interior plane rec_z.  ng=<n> records a
 * line (gxbeg + r*jgx, gybeg + r*jgy, gzbeg + r*jgz), r < n (the
 * reference's gxbeg..jgz, ng, ose_3d.cu:100-114; defaults jgx = 1,
 * jgy = jgz = 0); ngx/ngy record an ngx x ngy grid at depth gzbeg with
 * spacing jgx, jgy (default 1); csdgather=1 shifts every shot's receivers
 * by its source's offset from the model's first source.  Interior cells.
 * hilbert=unnorm|norm|analytic: Q's wavelet is nt H[s] (the reference),
 * H[s], or -H[s] (the analytic signal of the exp(-i L dt) propagator).
 * src_sigma=<cells>|auto replaces each point source by a normalised
 * Gaussian profile (ose_set_source_profile); auto takes sigma = 3.72 v dt /
 * (pi d) cells per axis, v the velocity at the source, so the profile's
 * spectrum is below 1e-3 at the temporal Nyquist pi / (v dt).
 * extent=exact (default; presets ref and c0) keeps each padded extent the
 * odd rule gives, which runs at any size up to 4096; extent=smooth (presets
 * c1-c4) widens an after-pad to the next extent with autotuned compiled
 * kernels, which changes the periodic domain and the taper: it reports each
 * change on stderr and in the info record ("extent_adjusted").            */
typedef struct oseh_sim oseh_sim;
int  oseh_sim_create(oseh_sim** out, const char* params);
/* Build velocities, pads, wavelets, orders and coefficients (host only).  */
int  oseh_sim_build(oseh_sim* sim);
/* oseh_sim_build, then register everything on a new ose_ctx
 * (ose_add_model / ose_add_shot).                                          */
int  oseh_sim_setup(oseh_sim* sim);
/* Run nsteps for every shot; ms receives the HIP-event time of the loop.  */
int  oseh_sim_run(oseh_sim* sim, int nsteps, double* ms);
ose_ctx* oseh_sim_ctx(oseh_sim* sim);
/* Boundary-network parameter file (.npy float32 array, or raw float32):
 * *n = count; values copied to out when out != NULL and cap >= count.
 * The simulation key bnet=<file> loads it for every model (guard=<g>,
 * bnet_taper=replace|keep select ose_set_boundary_net's options).          */
int oseh_read_params(const char* path, float* out, size_t cap, size_t* n);
/* JSON description: models (interior, padded, K, R, taper), shots,
 * interior and padded cells per step.                                      */
int  oseh_sim_info(oseh_sim* sim, char* buf, size_t cap);
/* Receivers of shot k (after oseh_sim_build): interior (x, y, z) triples,
 * the plane rec_z (x fastest) unless a line or grid is configured.  *n is
 * the capacity of xyz in receivers on entry (xyz may be NULL to query) and
 * the count on return; OSE_EINVAL if xyz is too small.                    */
int  oseh_sim_receivers(oseh_sim* sim, int shot, int* xyz, int* n);
/* Write traces (raw + .su) and, if requested, snapshots into dir.          */
int  oseh_sim_write(oseh_sim* sim, const char* dir);
/* Stream snapshots every `every` steps while the simulation runs
 * (ose_set_snapshots; 0 stops): keys snap_format=f32|bf16|q16|q8,
 * snap_region=padded|interior, snap_lanes=pq|p|q.  Per lane and shot,
 * float32 goes to dir/p-<it>-shot<k>.bin in the snapshot order below, the
 * encoded formats to dir/p-<it>-shot<k>.osnap (q- for Im u).  A failed
 * write is reported by the next oseh_sim_run that finds it, at the latest
 * by oseh_sim_write (which waits for every snapshot).                     */
int  oseh_sim_stream_snapshots(oseh_sim* sim, const char* dir, int every);
/* Write the current wavefield of every shot as p-<it>-shot<k>.bin (Re u)
 * and q-<it>-shot<k>.bin (Im u) in the snapshot order (ose_3d.cu:260-270). */
int  oseh_sim_snapshot(oseh_sim* sim, const char* dir);
/* Checkpoint / resume of a set-up simulation: the step counter, every
 * shot's wavefield and the traces recorded so far go to dir; restore reads
 * them into a simulation built with the same parameters and continues from
 * that step.  Long multi-shot runs restart without recomputing.           */
int  oseh_sim_checkpoint(oseh_sim* sim, const char* dir);
int  oseh_sim_restore(oseh_sim* sim, const char* dir);
const char* oseh_sim_error(oseh_sim* sim);
void oseh_sim_destroy(oseh_sim* sim);
/* Model / shot data of a simulation, for checking against the oracle.    */
int  oseh_sim_nmodels(oseh_sim* sim);
int  oseh_sim_model_dims(oseh_sim* sim, int model, int* dims /* 12: nx ny nz bl br bf bb bu bd nxb nyb nzb */);
int  oseh_sim_model_velocity(oseh_sim* sim, int model, float* v_pad);
int  oseh_sim_model_params(oseh_sim* sim, int model, double* R, int* K, double* coef /* K+1 */,
                           float* gx, float* gy, float* gz, int* rec_z);
int  oseh_sim_nshots(oseh_sim* sim);
int  oseh_sim_shot(oseh_sim* sim, int shot, int* model, int* src /* padded xyz */, int* nt,
                   float* s, float* h);

#ifdef __cplusplus
}
#endif
#endif /* OSE_HOST_H */
