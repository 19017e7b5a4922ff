/*
 * vfps_oracle.h -- CPU oracle for the per-step VFP phase-space update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU path of Inovesa (the non-OpenCL SourceMaps, PhaseSpace, ElectricField
 * and the impedance tables).  It exists to check the HIP library and to time
 * a CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * One source file is compiled twice:
 *   liboracle_f32.so  (real = float)   -- the parity target ("reference CPU path")
 *   liboracle_f64.so  (real = double)  -- the FP64 drift oracle
 *
 * Every function names the reference file:line it restates.  Parity status
 * of each function (what pins it) is listed in the tests that exercise it;
 * functions not pinned by a golden vector or closed form are marked
 * "parity unpinned" below.
 *
 * Floating-point contract: the library is compiled with -ffp-contract=off so
 * that every operation rounds as written in the reference source.
 */
#ifndef VFPS_ORACLE_H
#define VFPS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef VFPS_ORACLE_DOUBLE
typedef double real;
#else
typedef float real;
#endif

/* inc/SM/SourceMap.hpp:37-40  (struct hi {meshindex_t index; interpol_t weight;}) */
typedef struct {
    uint32_t index;
    real weight;
} or_hi;

typedef struct {
    real re;
    real im;
} or_cplx;

#ifdef __cplusplus
extern "C" {
#endif

/* size of one or_hi in bytes (lets the Python side pick the right dtype) */
size_t or_sizeof_hi(void);
size_t or_sizeof_real(void);

/* ---- inc/PS/Ruler.hpp:31-51 ------------------------------------------- */
void or_ruler(uint32_t steps, real min, real max,
              real* data, real* delta, real* zerobin);

/* ---- src/HelperFunctions.cpp:178-189 ---------------------------------- */
uint64_t or_upper_power_of_two(uint64_t v);

/* ---- src/PS/PhaseSpace.cpp:486-501 ------------------------------------ */
void or_simpson_weights(uint32_t n, real delta0, real* ws);

/* ---- src/PS/PhaseSpace.cpp:471-484 (gaus) ----------------------------- */
void or_gaus(uint32_t n, const real* axis, double zoom, real* out);

/* src/PS/PhaseSpace.cpp:17-70 + 457-469: the default-data constructor:
 * gaussians on both axes, outer product, projection, integral, normalise,
 * then projections and integral again.  data: [nb][n][n]. */
int or_phasespace_gaussian(uint32_t nb, uint32_t n,
                           real qmin, real qmax, real pmin, real pmax,
                           const real* filling_set, double zoom,
                           real* data, real* proj_x, real* proj_y,
                           real* filling, real* integral);

/* src/PS/PhaseSpace.cpp:47-51: filling must sum to 1 (rounded at 1e-5) */
int or_filling_ok(uint32_t nb, const real* filling_set);

/* ---- src/PS/PhaseSpace.cpp:313-321 (updateXProjection, CPU) ----------- */
void or_project_x(uint32_t nb, uint32_t n, const real* data, const real* ws,
                  real* proj);
/* ---- src/PS/PhaseSpace.cpp:325-340 (updateYProjection) ---------------- */
void or_project_y(uint32_t nb, uint32_t n, const real* data, const real* ws,
                  real* proj);
/* ---- src/PS/PhaseSpace.cpp:229-237 (integrate, CPU, real accumulator) - */
void or_integrate(uint32_t nb, uint32_t n, const real* proj, const real* ws,
                  real* filling, real* integral);
/* FP64 integral (the build computes the charge integral in FP64) */
void or_integrate_f64(uint32_t nb, uint32_t n, const real* proj,
                      const real* ws, double* filling, double* integral);
/* ---- src/PS/PhaseSpace.cpp:342-372 (normalize) ------------------------ */
void or_normalize(uint32_t nb, uint32_t n, real* data,
                  const real* filling_set, const real* filling);
/* ---- src/PS/PhaseSpace.cpp:241-289 (average + variance) --------------- */
void or_variance(uint32_t nb, uint32_t n, const real* proj_axis,
                 const real* axis, real delta, const real* filling_set,
                 const real* filling, real* mean, real* var, real* rms);

/* ---- src/SM/SourceMap.cpp:142-170 (calcCoefficiants) ------------------ */
void or_calc_coefficients(real* ic, real f, uint32_t it);

/* ---- src/SM/SourceMap.cpp:75-109 (generic apply) ---------------------- */
void or_sourcemap_apply(uint32_t ncells, uint32_t ip, const or_hi* hinfo,
                        const real* in, real* out);

/* ---- src/SM/KickMap.cpp:339-389 (updateSM) ---------------------------- */
void or_kick_update_sm(uint32_t kd, uint32_t noff, const real* offset,
                       uint32_t it, or_hi* hinfo);
/* ---- src/SM/KickMap.cpp:221-243 (apply, x kick) ----------------------- */
void or_kick_apply_x(uint32_t nb, uint32_t kd, uint32_t pd, uint32_t ip,
                     const or_hi* hinfo, const real* in, real* out);
/* ---- src/SM/KickMap.cpp:244-267 (apply, y kick) ----------------------- */
void or_kick_apply_y(uint32_t nb, uint32_t kd, uint32_t pd, uint32_t ip,
                     const or_hi* hinfo, const real* in, real* out);

/* ---- src/SM/DriftMap.cpp:11-38 (offsets) ------------------------------ */
void or_drift_offsets(uint32_t n, const real* axis1, real delta0,
                      const real* slip, uint32_t nslip, real pscale, real E0,
                      real* offset);
/* ---- src/SM/RFKickMap.cpp:15-83 (offsets, linear and sinusoidal) ------- */
void or_rfkick_offsets_linear(uint32_t n, real zerobin0, real angle,
                              real qscale, real f_RF, real delta0,
                              real* offset);
void or_rfkick_offsets(uint32_t n, const real* axis0, real revolutionpart,
                       real V_RF, real f_RF, real V0, real qscale,
                       real delta1, real pscale, real* offset);
/* _calcKick(phase, ampl) of src/SM/RFKickMap.cpp:59-75 (DynamicRFKickMap) */
void or_rfkick_offsets_at(uint32_t n, const real* axis0, real revolutionpart,
                          real V_RF, real f_RF, real V0, real qscale,
                          real delta1, real pscale, real phase, real ampl,
                          real* offset);
void or_rfkick_offsets_linear_at(uint32_t n, real zerobin0, real angle,
                                 real qscale, real f_RF, real delta0,
                                 real syncphase, real phase, real ampl,
                                 real* offset);

/* ---- src/SM/RotationMap.cpp:190-262 (genHInfo) ------------------------ */
void or_rotation_hinfo(uint32_t xsize, uint32_t ysize,
                       const real* axis0, real delta0, real zb0,
                       const real* axis1, real delta1, real zb1,
                       real cos_dt, real sin_dt, uint32_t it,
                       uint32_t x0, uint32_t y0, or_hi* myhinfo);
/* src/SM/RotationMap.cpp:12-95: constructor with rotmapsize = nxy.
 * Returns -1 (std::invalid_argument in the reference) when clamping is
 * requested without a cubic map. */
int or_rotation_build(uint32_t n, const real* axis0, real delta0, real zb0,
                      const real* axis1, real delta1, real zb1, real angle,
                      uint32_t it, int clamp, uint32_t rotmapsize,
                      or_hi* hinfo, real* cos_dt, real* sin_dt);
/* src/SM/RotationMap.cpp:97-176 (apply; rotmapsize = 0 regenerates taps) */
void or_rotation_apply(uint32_t n, uint32_t it, int clamp,
                       uint32_t rotmapsize, const or_hi* hinfo,
                       const real* axis0, real delta0, real zb0,
                       const real* axis1, real delta1, real zb1,
                       real cos_dt, real sin_dt, or_hi* scratch,
                       const real* in, real* out);

/* ---- src/SM/FokkerPlanckMap.cpp:11-118 (weights) ---------------------- *
 * deriv 3 = two_sided, 4 = cubic (reference), 5 = five-point (new, this
 * project's extension; the reference has no counterpart).  Returns the
 * number of taps per row (ip) or -1 for an unknown type. */
int or_fp_build(uint32_t n, uint32_t fptype, real e1, uint32_t deriv,
                const real* axis1, real delta1, real zb1, or_hi* hinfo);
/* ---- src/SM/FokkerPlanckMap.cpp:194-212 (apply) ----------------------- */
void or_fp_apply(uint32_t nb, uint32_t n, uint32_t ip, const or_hi* hinfo,
                 const real* in, real* out);

/* ---- FFT (stands in for FFTW's fftwf_plan_dft_r2c_1d / c2r_1d; computed in
 * double and rounded to real, the correctly rounded FP32 transform) ------ *
 * r2c: out[k] = sum_x in[x] exp(-2 pi i k x / n), k = 0..n/2.
 * c2r: out[x] = Re W0 + 2 sum_{k=1}^{n/2-1} Re(W_k e^{+2pi i k x/n})
 *               + Re W_{n/2} (-1)^x     (unnormalised, imag of DC and
 *               Nyquist ignored as FFTW's c2r does). */
void or_fft_r2c(uint32_t n, const real* in, or_cplx* out);
void or_fft_c2r(uint32_t n, const or_cplx* in, real* out);
/* direct O(n^2) DFT in double, for pinning the FFT on small n */
void or_dft_r2c_f64(uint32_t n, const double* in, double* out_re,
                    double* out_im);

/* ---- src/PS/ElectricField.cpp:255-263 (padBunchProfiles) -------------- */
void or_pad_profiles(uint32_t nb, uint32_t n, const real* proj,
                     uint32_t nmax, const uint32_t* buckets,
                     uint32_t spacing_bins, real* padded);
/* ---- src/PS/ElectricField.cpp:186-253 (wakePotential, CPU) ------------ *
 * Z: nmax entries.  wakescaling is the constructor argument (not yet
 * divided by nmax; the division of inc/PS/ElectricField.hpp:99-100 and
 * src/PS/ElectricField.cpp:58 happens here).  W[nmax/2] is forced to 0.
 * wp: [nb][n].  padded_out (nmax) may be NULL. */
void or_wake_potential(uint32_t nb, uint32_t n, const real* proj,
                       const or_cplx* Z, uint32_t nmax,
                       const uint32_t* buckets, uint32_t spacing_bins,
                       real wakescaling_arg, real* wp, real* padded_out);
/* ---- src/PS/ElectricField.cpp:133-184 (updateCSR, CPU) ---------------- *
 * spectrum: [nb][nmax], power: [nb]. */
void or_update_csr(uint32_t nb, uint32_t n, const real* proj,
                   const or_cplx* Z, uint32_t nmax, real delta0,
                   real qscale, real fc, real* spectrum, real* power);
/* inc/PS/ElectricField.hpp:81-103: wake scaling passed to the field */
real or_wakescaling(double Ib, double dt, double qscale, real delta1,
                    double sigma_delta, double E0);

/* ---- impedances (src/Z/ sources) ----------------------------------------- */
void or_z_const(uint32_t n, real re, real im, or_cplx* out);
void or_z_freespace(uint32_t n, real f_rev, real f_max, or_cplx* out);
void or_z_resistive_wall(uint32_t n, real f0, real f_max, double L,
                         double s, double xi, double b, or_cplx* out);
void or_z_collimator(uint32_t n, double outer, double inner, or_cplx* out);
/* ParallelPlatesCSR, FP64 Airy sums.  Returns the number of Airy terms
 * summed.  Parity unpinned at the Boost.Math overflow boundary: the sum
 * stops at the first odd p for which Bi'(u) or Bi(u) exceeds DBL_MAX. */
long or_z_parallel_plates(uint32_t n, real f0, real f_max, double g,
                          or_cplx* out);
/* Airy functions for x >= 0 in extended precision (Maclaurin series for
 * Bi, Bi'; K_{1/3}, K_{2/3} quadrature for Ai, Ai').  Results as double;
 * returns 1 if Bi or Bi' overflow double. */
int or_airy(double x, double* ai, double* aip, double* bi, double* bip);

/* ---- driver: one simulation, the loop of src/main.cpp:948-1053 -------- */
typedef struct or_sim or_sim;

typedef struct {
    uint32_t n;
    real qmin, qmax, pmin, pmax;
    real qscale;           /* "Meter" scale of axis 0 (bunch length) */
    real pscale;           /* "ElectronVolt" scale of axis 1 (dE) */
    uint32_t scheme;       /* 0: wake, RF, drift, FP (src/main.cpp:1034-1041)
                              1: wake, rotation, FP */
    uint32_t interp;       /* interpolation points (1..4) */
    int32_t rot_clamp;
    real angle;            /* rotation angle per step */
    int32_t rf_linear;
    real revolutionpart, V_RF, f_RF, V0;
    real slip[3];
    uint32_t nslip;
    real E0;
    int32_t use_fp;        /* 0: Identity (src/main.cpp:704-707) */
    uint32_t fptype;
    real e1;
    uint32_t deriv;
    int32_t use_wake;      /* 0: Identity instead of the wake map */
    uint32_t nmax;
    uint32_t spacing_bins;
    real wakescaling_arg;
    int32_t use_csr;       /* compute CSR power every step */
    real fc;
    uint32_t renormalize;  /* RenormalizeCharge: integrateAndNormalize on
                              every step whose count is a multiple
                              (src/main.cpp:953-955); 0: never */
} or_sim_cfg;

or_sim* or_sim_create(const or_sim_cfg* cfg, const real* init,
                      const or_cplx* Zwake, const or_cplx* Zcsr);
void or_sim_destroy(or_sim* s);
/* run nsteps; csr_power (nsteps, may be NULL) gets the CSR power at each
 * step (computed from the profile the wake of that step uses);
 * profiles (nsteps*n, may be NULL) the x projection after each step. */
void or_sim_run(or_sim* s, uint32_t nsteps, real* csr_power, real* profiles);
void or_sim_get(const or_sim* s, real* mesh, real* proj, real* wake);
/* one step split into its parts, for per-map parity */
void or_sim_set_mesh(or_sim* s, const real* mesh);

#ifdef __cplusplus
}
#endif

/* ---- passive particle tracking: SourceMap::applyTo (pos = {x, y}) ---- */
void or_track_kick(int axis, uint32_t pd, uint32_t kd, const real* offset,
                   real* pos);
void or_track_rotation(real cos_dt, real sin_dt, real zb0, real zb1,
                       real* pos);
void or_track_fp(int mode, uint32_t ip, uint32_t ysize, uint32_t nx,
                 const or_hi* hinfo, const real* data_in, real* pos);

#endif
