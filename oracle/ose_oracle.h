/*
 * ose_oracle.h -- FP64 CPU restatement of the 3D-OSE time stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libose_hip.so, libose_host.so, the ose3d driver)
 * never links or calls it.
 *
 * Every function restates one piece of the reference (Littleke/3D-OSE)
 * in double precision, following the reference's own operation order:
 *
 *   orc_fft*            cufftXtExecDescriptorC2C / cufftExecR2C,C2R
 *                       (ose_3d_lib.cu:668-820, ose_3d.cu:176-178):
 *                       own mixed-radix Stockham FFT, any length
 *   orc_ricker          generate_wavelet_gpu      ose_3d_kernel.cu:3-13
 *   orc_hilbert         Alloc_wavelet R2C->hilbert_1d->C2R
 *                       ose_3d_lib.cu:575-583, ose_3d_kernel.cu:154-172
 *   orc_bessel_jn       jnf() in step1to4_3d/add_bessel_3d
 *                       ose_3d_kernel.cu:310,388 (glibc jn, FP64)
 *   orc_taper_ref       generate_att               ose_3d_kernel.cu:242-261
 *   orc_judge_odd       judge_odd                  ose_3d_lib.cu:36-44
 *   orc_velocity_*      input_file_xyz_boundary + add_pml_layers_v_h
 *                       ose_3d_lib.cu:887-1032
 *   orc_apply_L         memcpy_between_device_cr -> FFT -> scale_kxkz_3D
 *                       -> IFFT -> scal_vr_3d    ose_3d_lib.cu:662-687
 *   orc_step            wavefield_extro_ori2       ose_3d_lib.cu:639-874
 *   orc_run             main() shot/time loop      ose_3d.cu:242-279
 *   orc_run_profile     the same loop with a spatially extended source
 *                       (not in the reference; a unit-vector profile is
 *                       orc_run's point source, pinned bit for bit in
 *                       tests/test_oracle.py, and superposition of two
 *                       point runs pins the sum over a profile)
 *
 * Arrays are [z][y][x] with x fastest (ose_3d_kernel.cu:35).
 */
#ifndef OSE_ORACLE_H
#define OSE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT ------------------------------------------------------------ */
/* Unnormalised in-place DFT of interleaved complex doubles (re,im pairs).
 * sign = -1 forward (e^{-2pi i nk/N}), +1 inverse.                       */
int  orc_fft1d(double* a, int n, int sign);
/* 3D transform of a [nz][ny][nx] complex array.  Threads: OpenMP.       */
int  orc_fft3d(double* a, int nx, int ny, int nz, int sign);
/* O(N^2) brute-force 3D DFT; test helper for tiny grids only.          */
void orc_dft3d_brute(const double* in, double* out, int nx, int ny, int nz, int sign);

/* ---- source wavelet -------------------------------------------------- */
/* Ricker delayed by `delay` samples: tmp = pi*fm*((i-delay)*dt - 1/fm),
 * w = amp*(1-2 tmp^2) exp(-tmp^2).  Reference: delay = 40.            */
void orc_ricker(double* w, int nt, double fm, double dt, double amp, int delay);
/* Reference per-shot Q source: C2R( sgn-rotated R2C(s) ) without the 1/nt
 * normalisation, i.e. nt times the discrete Hilbert transform.          */
void orc_hilbert(const double* s, double* h, int nt);

/* ---- Bessel / taper / grid ----------------------------------------- */
double orc_bessel_jn(int n, double x);
void   orc_taper_ref(double* a, int before, int nn, int after);
void   orc_taper_cerjan(double* a, int before, int nn, int after, double alpha);
int    orc_judge_odd(int* after, int* nb);

/* Read reference file order [x][y][z] (z fastest) into padded [z][y][x]
 * and edge-replicate into the pads.                                     */
void orc_velocity_from_file_order(const float* file_xyz, int nx, int ny, int nz,
                                  int bl, int bf, int bu, int nxb, int nyb, int nzb,
                                  double* v_zyx);
/* Edge-replicate an interior [nz][ny][nx] cube into padded [z][y][x].   */
void orc_velocity_pad_interior(const float* v_int_zyx, int nx, int ny, int nz,
                               int bl, int bf, int bu, int nxb, int nyb, int nzb,
                               double* v_zyx);

/* ---- the time stepper ------------------------------------------------ */
typedef struct {
    int nxb, nyb, nzb;          /* padded extents                        */
    double dx, dy, dz, dt;
    double R;                   /* spectral radius bound                 */
    int K;                      /* highest order = operator applications */
    const double* coef;         /* coef[0..K] = J0, 2J1, ..., 2JK         */
    const double* v;            /* padded velocity [z][y][x]             */
    const double* gx;           /* taper profiles (nxb, nyb, nzb)        */
    const double* gy;
    const double* gz;
} orc_model;

/* out = (v/R) * IFFT3(|k| FFT3(in)) / N for a complex field (the operator
 * is real, so Re and Im are transformed independently).                */
int orc_apply_L(const orc_model* m, const double* in_c, double* out_c);

/* One OSE step on (P, Q) in place, reference lane formulation.          */
int orc_step(const orc_model* m, double* P, double* Q);

/* Full single-shot run: inject s/h at padded linear index src, step,
 * record Q on padded depth plane rz over interior x in [bl,bl+nx),
 * y in [bf,bf+ny).  traces: [nrec][nt], nrec = nx*ny (reference order
 * ig = iy*nx + ix).  P, Q: state in/out (nxb*nyb*nzb each).          */
int orc_run(const orc_model* m, int nt, long src, const double* s, const double* h,
            int rz, int bl, int bf, int nx, int ny, double* traces,
            double* P, double* Q);

/* The same run with a spatially extended source: before each step
 * P[z][y][x] += s[it] px[x] py[y] pz[z] and Q likewise with h (px, py, pz:
 * full-length profiles, zero off the support).  Not in the reference, whose
 * source is one cell (Xt_add_wavelet_sxyz, ose_3d_lib.cu:610-633): a point
 * injected every dt excites the grid's modes above the temporal Nyquist
 * pi/dt, a profile band-limited below it does not.  A unit-vector profile
 * is orc_run's point source.                                            */
int orc_run_profile(const orc_model* m, int nt, const double* px, const double* py, const double* pz,
                    const double* s, const double* h, int rz, int bl, int bf, int nx, int ny,
                    double* traces, double* P, double* Q);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
