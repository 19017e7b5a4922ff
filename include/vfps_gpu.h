/*
 * vfps_gpu.h -- C ABI of the MI355X (gfx950) VFP phase-space step.
 *
 * The library replaces the device back end of Inovesa's per-step update:
 * the OpenCL handle `oclhptr_t` (inc/CL/OpenCLHandler.hpp:36-41) that every
 * PhaseSpace, SourceMap, ElectricField and Impedance constructor receives.
 * Each entry point below names the reference interface it stands for.
 *
 * Conventions
 *   - Every call returns 0 on success and a negative vfps_status on failure;
 *     vfps_last_error() gives a message (thread-local).  No C++ exception
 *     crosses this boundary.
 *   - One HIP stream per device handle; calls are asynchronous on it, except
 *     those that return data to host memory, which synchronise.
 *   - Host pointers are borrowed for the duration of the call only.  Device
 *     memory is owned by the handles.
 *   - Meshes are FP32, layout [bunch][x][y] row-major, y contiguous
 *     (inc/PS/PhaseSpace.hpp:444-447).
 */
#ifndef VFPS_GPU_H
#define VFPS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFPS_API __attribute__((visibility("default")))

typedef enum {
    VFPS_OK = 0,
    VFPS_ERR_INVALID = -1,  /* bad argument (std::invalid_argument in the reference) */
    VFPS_ERR_HIP = -2,      /* HIP runtime error */
    VFPS_ERR_NOMEM = -3,
    VFPS_ERR_UNSUPPORTED = -4
} vfps_status;

typedef struct vfps_dev vfps_dev;
typedef struct vfps_mesh vfps_mesh;
typedef struct vfps_map vfps_map;
typedef struct vfps_impedance vfps_impedance;
typedef struct vfps_field vfps_field;
typedef struct vfps_step vfps_step;
typedef struct vfps_batch vfps_batch;
typedef struct vfps_mlp vfps_mlp;

typedef struct {
    float re;
    float im;
} vfps_cplx;

/* ---- errors, version ---------------------------------------------------- */
VFPS_API const char* vfps_last_error(void);
VFPS_API const char* vfps_version(void);

/* ---- device (replaces OCLH, inc/CL/OpenCLHandler.hpp:63-314) ------------ */
VFPS_API int vfps_dev_open(int ordinal, vfps_dev** out);
VFPS_API int vfps_dev_close(vfps_dev* dev);
/* OCLH::flush/finish (src/main.cpp:1046-1050) */
VFPS_API int vfps_sync(vfps_dev* dev);

/* ---- PhaseSpace (inc/PS/PhaseSpace.hpp:40-552) -------------------------- */
/* PhaseSpace(qmin,qmax,qscale,pmin,pmax,pscale,oclh,Qb,Ib,filling,zoom,data)
 * (src/PS/PhaseSpace.cpp:149-168); nx = ny = n, nb = number of bunches.
 * data == NULL builds the Gaussian of src/PS/PhaseSpace.cpp:56-63 on the
 * device; otherwise the mesh is copied from data ([nb][n][n]).
 * Fails with VFPS_ERR_INVALID if filling does not sum to 1 (:47-51). */
VFPS_API int vfps_mesh_create(vfps_dev* dev, uint32_t nb, uint32_t n,
                              float qmin, float qmax, double qscale,
                              float pmin, float pmax, double pscale,
                              double charge, double current,
                              const float* filling, double zoom,
                              const float* data, vfps_mesh** out);
/* copy constructor (src/PS/PhaseSpace.cpp:170-180) */
VFPS_API int vfps_mesh_clone(const vfps_mesh* src, vfps_mesh** out);
VFPS_API int vfps_mesh_destroy(vfps_mesh* m);
/* syncCLMem(cpu2dev) / syncCLMem(dev2cpu) (src/PS/PhaseSpace.cpp:381-414) */
VFPS_API int vfps_mesh_upload(vfps_mesh* m, const float* data);
VFPS_API int vfps_mesh_download(vfps_mesh* m, float* data);
/* operator= / swap (src/PS/PhaseSpace.cpp:374-379, 503-506) */
VFPS_API int vfps_mesh_copy(vfps_mesh* dst, const vfps_mesh* src);
/* getData(): device pointer, valid while the mesh lives */
VFPS_API int vfps_mesh_device_ptr(vfps_mesh* m, float** dptr);
/* updateXProjection (src/PS/PhaseSpace.cpp:291-323) */
VFPS_API int vfps_mesh_project_x(vfps_mesh* m);
/* updateYProjection (src/PS/PhaseSpace.cpp:325-340) */
VFPS_API int vfps_mesh_project_y(vfps_mesh* m);
/* getProjection(axis) ([nb][n]); synchronises */
VFPS_API int vfps_mesh_get_projection(vfps_mesh* m, int axis, float* out);
/* setProjection(axis, bunch, ...) (inc/PS/PhaseSpace.hpp:197-207) */
VFPS_API int vfps_mesh_set_projection(vfps_mesh* m, int axis,
                                      const float* proj);
/* integrate (src/PS/PhaseSpace.cpp:200-239), accumulated in FP64.
 * filling (nb) and integral may be NULL; non-NULL outputs synchronise. */
VFPS_API int vfps_mesh_integrate(vfps_mesh* m, double* filling,
                                 double* integral);
/* normalize (src/PS/PhaseSpace.cpp:342-372) with the last integrate() */
VFPS_API int vfps_mesh_normalize(vfps_mesh* m);
/* variance(axis) (src/PS/PhaseSpace.cpp:241-289): mean, variance and rms
 * per bunch from the current projection of that axis; synchronises */
VFPS_API int vfps_mesh_variance(vfps_mesh* m, int axis, float* mean,
                                float* var, float* rms);
/* geometry: getDelta(axis), zerobin, Simpson weights (n) */
VFPS_API int vfps_mesh_info(const vfps_mesh* m, uint32_t* nb, uint32_t* n,
                            float* delta0, float* delta1, float* zerobin0,
                            float* zerobin1);
VFPS_API int vfps_mesh_axis(const vfps_mesh* m, int axis, float* out);
VFPS_API int vfps_mesh_simpson_weights(const vfps_mesh* m, float* out);

/* ---- SourceMaps (inc/SM/SourceMap.hpp:53-87) ---------------------------- */
/* SourceMap::InterpolationType (inc/SM/SourceMap.hpp:22-27) */
enum { VFPS_INTERP_NONE = 1, VFPS_INTERP_LINEAR = 2,
       VFPS_INTERP_QUADRATIC = 3, VFPS_INTERP_CUBIC = 4 };
/* KickMap(in,out,it,interpol_clamp,kd,oclh) (src/SM/KickMap.cpp:25-192);
 * axis 0 = Axis::x (displacement along x, per y), 1 = Axis::y.
 * The reference CPU path ignores interpol_clamp for kick maps; so does
 * this one. */
VFPS_API int vfps_map_create_kick(vfps_mesh* in, vfps_mesh* out, int interp,
                                  int clamp, int axis, vfps_map** map);
/* KickMap::swapOffset (inc/SM/KickMap.hpp:51-55): offsets in mesh units,
 * n entries (pd * nb) */
VFPS_API int vfps_map_set_offset(vfps_map* map, const float* offset,
                                 size_t n);
/* KickMap::getForce (inc/SM/KickMap.hpp:47-49); synchronises */
VFPS_API int vfps_map_get_offset(vfps_map* map, float* offset, size_t n);
/* DriftMap(in,out,slip,E0,it,clamp,oclh) (src/SM/DriftMap.cpp:11-38) */
VFPS_API int vfps_map_create_drift(vfps_mesh* in, vfps_mesh* out,
                                   const float* slip, uint32_t nslip,
                                   float E0, int interp, int clamp,
                                   vfps_map** map);
/* RFKickMap linear (src/SM/RFKickMap.cpp:15-34) */
VFPS_API int vfps_map_create_rfkick_linear(vfps_mesh* in, vfps_mesh* out,
                                           float angle, float f_RF,
                                           int interp, int clamp,
                                           vfps_map** map);
/* RFKickMap sinusoidal (src/SM/RFKickMap.cpp:36-57) */
VFPS_API int vfps_map_create_rfkick(vfps_mesh* in, vfps_mesh* out,
                                    float revolutionpart, float V_RF,
                                    float f_RF, float V0, int interp,
                                    int clamp, vfps_map** map);
/* DynamicRFKickMap, sinusoidal (src/SM/DynamicRFKickMap.cpp:45-67):
 * an RFKickMap whose kick is recomputed before every apply() from the next
 * entry of a modulation queue of `steps` (phase, amplitude) pairs built at
 * construction (__calcModulation, :80-95: normally distributed phase and
 * amplitude noise of phasespread/sqrt(revolutionpart) and
 * amplspread/sqrt(revolutionpart), plus a phase modulation
 * modampl*sin(2 pi modtimeincrement i)).  The noise comes from
 * std::mt19937 seeded with `seed` (the reference seeds it from
 * std::random_device) and std::normal_distribution<float>.
 * revolutionpart is the reference's double (the kick holds it as float).
 * An apply() beyond `steps` fails with VFPS_ERR_INVALID. */
VFPS_API int vfps_map_create_dynrfkick(vfps_mesh* in, vfps_mesh* out,
                                       double revolutionpart, float V_RF,
                                       float f_RF, float V0,
                                       float phasespread, float amplspread,
                                       float modampl, double modtimeincrement,
                                       uint32_t steps, uint32_t seed,
                                       int interp, int clamp, vfps_map** map);
/* DynamicRFKickMap, linear (src/SM/DynamicRFKickMap.cpp:15-38): the
 * linear RF kick of src/SM/RFKickMap.cpp:59-69 with the same modulation
 * (synchronous phase 0) */
VFPS_API int vfps_map_create_dynrfkick_linear(vfps_mesh* in, vfps_mesh* out,
                                              float angle,
                                              double revolutionpart,
                                              float f_RF, float phasespread,
                                              float amplspread, float modampl,
                                              double modtimeincrement,
                                              uint32_t steps, uint32_t seed,
                                              int interp, int clamp,
                                              vfps_map** map);
/* DynamicRFKickMap::getPastModulation (src/SM/DynamicRFKickMap.cpp:113-119):
 * entries [first, first+count) of the modulation queue ([count][2]: phase,
 * amplitude) and the number applied so far */
VFPS_API int vfps_map_get_modulation(const vfps_map* map, uint32_t first,
                                     uint32_t count, float* out,
                                     uint32_t* applied);
/* the modulation queue alone (host only, no device): [steps][2] */
VFPS_API int vfps_dynrf_modulation(uint32_t steps, uint32_t seed,
                                   float syncphase, double revolutionpart,
                                   float phasespread, float amplspread,
                                   float modampl, double modtimeincrement,
                                   float* out);
/* RotationMap(in,out,xsize,ysize,angle,it,clamped,rotmapsize,oclh)
 * (src/SM/RotationMap.cpp:12-95).  The taps are computed on the fly for
 * any rotmapsize; rotmapsize only selects the reference semantics (clamping
 * needs cubic interpolation and rotmapsize > 0, else VFPS_ERR_INVALID). */
VFPS_API int vfps_map_create_rotation(vfps_mesh* in, vfps_mesh* out,
                                      float angle, int interp, int clamp,
                                      uint32_t rotmapsize, vfps_map** map);
/* FokkerPlanckMap(in,out,xsize,ysize,fptype,fptrack,e1,dt,oclh)
 * (src/SM/FokkerPlanckMap.cpp:11-118).  fptype: 0 none, 1 damping only,
 * 2 diffusion only, 3 full; deriv: 3 two-sided, 4 cubic, 5 five-point
 * (this library's extension, no reference counterpart). */
VFPS_API int vfps_map_create_fokkerplanck(vfps_mesh* in, vfps_mesh* out,
                                          int fptype, int fptrack, float e1,
                                          int deriv, vfps_map** map);
/* Identity (inc/SM/Identity.hpp:14-61) */
VFPS_API int vfps_map_create_identity(vfps_mesh* in, vfps_mesh* out,
                                      vfps_map** map);
/* WakePotentialMap(in,out,field,it,clamp,oclh)
 * (src/SM/WakePotentialMap.cpp:11-25) */
VFPS_API int vfps_map_create_wakepotential(vfps_mesh* in, vfps_mesh* out,
                                           vfps_field* field, int interp,
                                           int clamp, vfps_map** map);
/* WakePotentialMap::update (src/SM/WakePotentialMap.cpp:34-50): wake
 * potential of the field's phase space becomes this map's offset, on the
 * device */
VFPS_API int vfps_map_update(vfps_map* map);
/* SourceMap::applyToAll (src/SM/SourceMap.cpp:111-117): the map's applyTo
 * (KickMap src/SM/KickMap.cpp:280-304, RotationMap
 * src/SM/RotationMap.cpp:178-188, FokkerPlanckMap with the FPTracking the
 * map was built with, src/SM/FokkerPlanckMap.cpp:214-268; Identity moves
 * nothing) on n particles in grid coordinates ([n][2]: x, y; host array,
 * updated in place), on the device; synchronises.  The stochastic
 * Fokker-Planck tracking draws its noise from a counter-based generator
 * (seed, particle, call) where the reference uses std::mt19937 seeded from
 * std::random_device. */
VFPS_API int vfps_map_apply_to(vfps_map* map, float* pos, uint32_t n,
                               uint64_t seed);
/* SourceMap::apply (src/SM/SourceMap.cpp:75-109 and overrides) */
VFPS_API int vfps_map_apply(vfps_map* map);
VFPS_API int vfps_map_destroy(vfps_map* map);

/* ---- Impedance (inc/Z/Impedance.hpp; src/Z/) --------------------------- */
/* Impedance(nfreqs, f_max): a zero table (src/Z/Impedance.cpp:48-57) */
VFPS_API int vfps_impedance_create(vfps_dev* dev, uint32_t nfreqs,
                                   float f_max, vfps_impedance** out);
/* Impedance(z, f_max) from a table (src/Z/Impedance.cpp:35-46); also the
 * result of Impedance(datafile, ...) once the file is read */
VFPS_API int vfps_impedance_create_table(vfps_dev* dev, const vfps_cplx* z,
                                         uint32_t nfreqs, float f_max,
                                         vfps_impedance** out);
VFPS_API int vfps_impedance_destroy(vfps_impedance* z);
/* operator+= of the named contribution, each computed on the device in
 * FP64 and rounded to complex<float>:
 *   FreeSpaceCSR(n, f_rev, f_max)          src/Z/FreeSpaceCSR.cpp:19-42
 *   ParallelPlatesCSR(n, f0, f_max, g)     src/Z/ParallelPlatesCSR.cpp:19-69
 *   ResistiveWall(n, f0, f_max, L, s, xi, b) src/Z/ResistiveWall.cpp:25-55
 *   ConstImpedance(n, f_max, Z)            src/Z/ConstImpedance.cpp:11-27
 *   CollimatorImpedance(n, f_max, outer, inner) src/Z/CollimatorImpedance.cpp */
VFPS_API int vfps_impedance_add_freespace_csr(vfps_impedance* z, float f_rev,
                                              float f_max);
VFPS_API int vfps_impedance_add_parallel_plates(vfps_impedance* z, float f0,
                                                float f_max, double g);
VFPS_API int vfps_impedance_add_resistive_wall(vfps_impedance* z, float f0,
                                               float f_max, double L,
                                               double s, double xi, double b);
VFPS_API int vfps_impedance_add_const(vfps_impedance* z, float re, float im);
VFPS_API int vfps_impedance_add_collimator(vfps_impedance* z, double outer,
                                           double inner);
VFPS_API int vfps_impedance_add(vfps_impedance* z, const vfps_impedance* rhs);
/* operator[] over the whole table; synchronises */
VFPS_API int vfps_impedance_download(vfps_impedance* z, vfps_cplx* out);
VFPS_API int vfps_impedance_nfreqs(const vfps_impedance* z, uint32_t* n);
/* FP64 Airy functions Ai, Ai', Bi, Bi' on the device for x >= 0 (the
 * kernel the ParallelPlatesCSR table uses); out: 4*n doubles, ovf: n flags
 * set where Bi or Bi' exceed the double range; synchronises */
VFPS_API int vfps_airy(vfps_dev* dev, const double* x, uint32_t n,
                       double* out, int32_t* ovf);

/* ---- ElectricField (inc/PS/ElectricField.hpp) --------------------------- */
/* ElectricField(ps, impedance, buckets, spacing_bins, oclh, f_rev,
 * revolutionpart, wakescaling) (src/PS/ElectricField.cpp:14-101).
 * wakescaling is the constructor argument; the field divides it by nmax
 * (inc/PS/ElectricField.hpp:99-100).  Pass 0 for a radiation-only field. */
VFPS_API int vfps_field_create(vfps_mesh* ps, vfps_impedance* z,
                               const uint32_t* buckets, uint32_t nbuckets,
                               uint32_t spacing_bins, double f_rev,
                               float revolutionpart, float wakescaling,
                               vfps_field** out);
/* the scaling argument of the second constructor
 * (inc/PS/ElectricField.hpp:81-103): Ib*dt*c/Meter/(dp*sigma_delta*E0) */
VFPS_API float vfps_field_wakescaling(double Ib, double dt, double qscale,
                                      float delta1, double sigma_delta,
                                      double E0);
VFPS_API int vfps_field_destroy(vfps_field* f);
/* padBunchProfiles (src/PS/ElectricField.cpp:255-263) */
VFPS_API int vfps_field_pad(vfps_field* f);
/* getPaddedBunchProfiles: nmax floats; synchronises */
VFPS_API int vfps_field_get_padded(vfps_field* f, float* out);
/* getPaddedWakePotential (inc/PS/ElectricField.hpp:170-171): the
 * unscaled inverse transform of the last wakePotential(), nmax floats;
 * synchronises */
VFPS_API int vfps_field_get_padded_wake(vfps_field* f, float* out);
/* wakePotential (src/PS/ElectricField.cpp:186-253); wp ([nb][n]) may be
 * NULL (the result stays on the device); non-NULL synchronises */
VFPS_API int vfps_field_wake_potential(vfps_field* f, float* wp);
/* No reference counterpart.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vfps_field_wake_potential without a field object, on
 * caller-owned device buffers.  proj: device [nb][n] x projections; z:
 * device impedance table of nmax entries; buckets: host array of nb bucket
 * numbers (profile b sits at buckets[b] * spacing_bins of the padded
 * array); wakescaling: the ElectricField constructor argument (divided by
 * nmax here, as there); wake: device [nb][n] output.  The library allocates
 * nothing: its scratch (FFT twiddle tables, bucket numbers) is a fixed
 * device array of 128.25 KiB per GPU, held until the process ends, so
 * nmax <= 8192 and nb <= 64 (larger: VFPS_ERR_INVALID).  Asynchronous on
 * the device stream.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other threads, other device handles of the same GPU)
 * race. */
VFPS_API int vfps_wake_potential_noalloc(vfps_dev* dev, const float* proj,
                                         uint32_t nb, uint32_t n,
                                         const vfps_cplx* z, uint32_t nmax,
                                         const uint32_t* buckets,
                                         uint32_t spacing_bins,
                                         float wakescaling, float* wake);
/* getWakeScaling (inc/PS/ElectricField.hpp:306-307) */
VFPS_API int vfps_field_get_wakescaling(const vfps_field* f, float* s);
/* updateCSR(fc) (src/PS/ElectricField.cpp:133-184); spectrum ([nb][nmax])
 * and power (nb) may be NULL; non-NULL synchronises */
VFPS_API int vfps_field_update_csr(vfps_field* f, float fc, float* power,
                                   float* spectrum);

/* ---- ring parameters (src/main.cpp:179-355) ----------------------------- */
typedef struct {
    uint32_t grid_size;       /* GridSize */
    double pq_size;           /* PhaseSpaceSize */
    double shift_x, shift_y;  /* PhaseSpaceShiftX/Y */
    double energy_spread;     /* sE */
    double beam_energy;       /* E0 (eV) */
    double bending_radius;    /* R; <= 0: from f_rev */
    double f_rev;             /* RevolutionFrequency */
    double harmonic_number;
    double rf_voltage;        /* V_RF */
    double f_s;               /* SyncFreq; 0: from alpha0 */
    double alpha0, alpha1, alpha2;
    double vacuum_gap;        /* VacuumChamberGap (m), <0 free space */
    double damping_time;      /* <0: from ring parameters */
    double padding;
    int32_t round_padding;
    uint32_t steps_per_ts;
    double rotations;
    double bunch_current;     /* single bunch */
    double cutoff_frequency;  /* fc */
    double steps_per_trev;    /* StepsPerRevolution; > 0 overrides
                                 steps_per_ts (src/main.cpp:284-286) */
} vfps_ring_params;

typedef struct {
    float qmin, qmax, pmin, pmax;
    double qscale;            /* bunch length bl (m): "Meter" scale */
    double pscale;            /* dE (eV): "ElectronVolt" scale */
    double R_bend, f0;        /* f0 = c/(2 pi R) */
    double V0, V_eff, f_s, alpha0;
    double f_RF;
    double steps, dt, revolutionpart;
    float angle;              /* 2 pi / steps */
    float slip[3];
    float e1;                 /* 2/(fs t_damp steps) */
    double t_damp;
    uint32_t padded_bins;     /* FFT length nmax */
    uint32_t spacing_bins;
    float fmax;
    double Ib;
    float wakescaling;        /* ElectricField scaling argument */
    double shield;            /* shielding parameter, g = gap */
    double Inorm;             /* scaling-law normalisation current */
    double Ith;               /* scaling-law threshold Inorm*(0.5+0.34*shield) */
    uint32_t laststep;
} vfps_derived;

VFPS_API int vfps_derive(const vfps_ring_params* p, vfps_derived* d);

/* ---- whole step on one mesh (src/main.cpp:948-1053) --------------------- *
 * A step object owns three meshes (grid_t1..t3), the maps and the fields;
 * run() advances nsteps steps with the wake of each step taken from the
 * profile of the previous one, exactly like the reference loop. */
typedef struct {
    uint32_t n;
    float qmin, qmax, pmin, pmax;
    double qscale, pscale;
    uint32_t scheme;          /* 0: wake, RF kick, drift, FP; 1: wake, rotation, FP */
    uint32_t interp;          /* interpolation points (1..4) */
    int32_t rot_clamp;
    float angle;              /* scheme 1: rotation angle per step */
    int32_t rf_linear;        /* scheme 0: RFKickMap flavour */
    float rf_angle;           /* linear RF kick angle */
    float revolutionpart, V_RF, f_RF, V0;
    float slip[3];
    uint32_t nslip;
    float E0;
    int32_t use_fp;
    uint32_t fptype;
    float e1;
    uint32_t deriv;
    int32_t use_wake;
    uint32_t spacing_bins;
    float wakescaling;        /* ElectricField scaling argument */
    int32_t use_csr;
    float fc;
    int32_t fused;            /* 1: fused kernels (default), 0: one kernel per map */
    uint32_t tile_rows;       /* 0: autotuned; rows per workgroup of the fused kernel */
    int32_t fast_projection;  /* 0: x projection summed in the reference's
                                 order (bit-exact); 1: wavefront tree sum */
    uint32_t wake_method;     /* fused step: 0 default (2), 1 FFT field kernel
                                 then the maps kernel, 2 direct: one kernel
                                 per step, wake and CSR power as FP64
                                 Toeplitz products of the profile */
    uint32_t threads;         /* fused step workgroup size (0: default) */
    int32_t no_graph;         /* 1: vfps_step_run launches every step itself
                                 instead of replaying a captured hipGraph */
    uint32_t renormalize;     /* RenormalizeCharge (src/main.cpp:953-955):
                                 > 0 renormalises the charge to the set
                                 bunch population before the maps of every
                                 step whose count is a multiple of it (the
                                 fused direct form only) */
} vfps_step_desc;

/* The FP64 Toeplitz kernels of the direct wake and CSR forms (host only,
 * no device needed).  With N = nmax, M = N/2, theta = 2 pi k d / N:
 *   g[d] = Re z_0 + 2 sum_{k=1}^{M-1} (Re z_k cos theta - Im z_k sin theta),
 * the C2R of z with its Nyquist bin zero, so that the wake potential of
 * src/PS/ElectricField.cpp:215-241 before scaling is sum_x' g[x-x'] p[x'];
 *   h[d] = sum_{k=0}^{M} fdelta renorm_k Re zc_k cos theta,
 * so that the CSR power of src/PS/ElectricField.cpp:133-184 is
 * sum_{x,x'} p[x] h[x-x'] p[x'].  g, h: 2n-1 doubles, index d + n - 1;
 * either may be NULL. */
VFPS_API int vfps_toeplitz_kernels(uint32_t nmax, uint32_t n,
                                   const vfps_cplx* z, const vfps_cplx* zc,
                                   const float* renorm, float fdelta,
                                   double* g, double* h);

VFPS_API int vfps_step_create(vfps_dev* dev, const vfps_step_desc* desc,
                              const float* init, vfps_impedance* zwake,
                              vfps_impedance* zcsr, vfps_step** out);
VFPS_API int vfps_step_destroy(vfps_step* s);
/* Advance nsteps (src/main.cpp:948-1053, nsteps passes of the loop body).
 * Unless no_graph is set, the steps replay captured hipGraphs: runs of 64
 * steps, then one run of the remaining nsteps % 64, each captured on first
 * use and kept per (length, starting parity).  Per-step CSR power and
 * bunch profiles are kept on the device (vfps_step_get_history). */
VFPS_API int vfps_step_run(vfps_step* s, uint32_t nsteps);
/* No reference counterpart: capture, without running them, the graphs
 * that vfps_step_run(s, nsteps) from the current state replays, so that a
 * timed run spends no time capturing. */
VFPS_API int vfps_step_prepare(vfps_step* s, uint32_t nsteps);
/* record per-step outputs of the next run() calls: capacity steps of CSR
 * power (and profiles if want_profiles) */
VFPS_API int vfps_step_record(vfps_step* s, uint32_t capacity,
                              int want_profiles);
VFPS_API int vfps_step_get_history(vfps_step* s, uint32_t* count,
                                   float* csr_power, float* profiles);
VFPS_API int vfps_step_get_mesh(vfps_step* s, float* mesh, float* proj,
                                float* wake);
VFPS_API int vfps_step_set_mesh(vfps_step* s, const float* mesh);
/* Particles tracked through every following step (src/main.cpp:1034-1041:
 * applyToAll after each map's apply), n positions in grid coordinates
 * ([n][2]), FokkerPlanckMap tracking fptrack (FPTracking 0-3; 2 needs the
 * per-map step, fused = 0), noise seed; n = 0 stops tracking.  The tracking
 * runs on the device beside the step (inside its captured graphs). */
VFPS_API int vfps_step_set_particles(vfps_step* s, const float* pos,
                                     uint32_t n, int fptrack, uint64_t seed);
/* the tracked positions ([n][2], n as set); synchronises */
VFPS_API int vfps_step_get_particles(vfps_step* s, float* pos, uint32_t n);
/* The outstep moments of the current state (src/main.cpp:961-967:
 * integrate(), variance(0), updateYProjection(), variance(1)), as
 * HDF5File::append reads them (src/IO/HDF5File.cpp:402-418): x projection
 * and y projection (n each), bunch population (the FP64 integral),
 * position getMoment(0,0), length getBunchLength(), energy average
 * getMoment(1,0), spread getEnergySpread(); any may be NULL; synchronises */
VFPS_API int vfps_step_moments(vfps_step* s, float* xproj, float* yproj,
                               float* population, float* position,
                               float* length, float* average, float* spread);
/* Advance nsteps and time the kernels with HIP events on the step's stream
 * (milliseconds per step): the direct form launches its one kernel nsteps
 * times back to back between two events (ms_maps; ms_wake = 0); the FFT
 * form brackets the field kernel (ms_wake) and the step kernel (ms_maps) of
 * each step. */
VFPS_API int vfps_step_timing(vfps_step* s, uint32_t nsteps, float* ms_wake,
                              float* ms_maps, float* ms_step);
/* one fused step with device timestamps at the phase boundaries of the
 * first workgroup of each kernel; ns_out (>= 16): [1..7] step kernel
 * (loads, wake products, wake kick, RF kick, drift/rotation, FP,
 * projection) and [9..13] field kernel (loads, forward FFT, spectrum pass,
 * inverse FFT, output), nanoseconds since the kernel's first stamp; [15]
 * the clock period; (nout >= 24) [16..21] step workgroups: first entry
 * after the field kernel, entry spread, shortest, median and longest
 * workgroup, last exit; (nout >= 24 + 2 tiles) [24 + 2t], [25 + 2t] entry
 * and exit of tile t after the first entry */
VFPS_API int vfps_step_phases(vfps_step* s, double* ns_out, uint32_t nout);
/* the configuration in use: out[0] fused, [1] wake method (1 FFT, 2 direct),
 * [2] tile rows, [3] threads, [4] LDS buffer rows, [5] LDS bytes (nout >= 6),
 * [6] persistent launch in a batch (nout >= 7) */
VFPS_API int vfps_step_config(vfps_step* s, uint32_t* out, uint32_t nout);

/* ---- batch of independent meshes ------------------------------------- *
 * No reference counterpart: a scan over currents or chamber gaps runs the
 * time loop of src/main.cpp:948-1053 once per point.  Here ninst instances
 * of any mesh sizes (each described as for vfps_step_create, with its own
 * initial mesh and impedances) advance together: one step kernel launch per
 * step covers every row tile of every instance (one launch per kernel
 * variant and workgroup size), captured into a hipGraph.  An instance's
 * wake_method 1 selects the FFT wake (one field-kernel launch per step
 * covers every such instance, a workgroup each), anything else the direct
 * wake.  Each instance equals a fused single-mesh step with the same wake
 * form bit for bit. */
VFPS_API int vfps_batch_create(vfps_dev* dev, uint32_t ninst,
                               const vfps_step_desc* descs,
                               const float* const* inits,
                               vfps_impedance* const* zwake,
                               vfps_impedance* const* zcsr, vfps_batch** out);
VFPS_API int vfps_batch_destroy(vfps_batch* b);
/* nsteps steps of every instance, replayed from captured graphs as in
 * vfps_step_run (launched from the host when the first instance's
 * description sets no_graph) */
VFPS_API int vfps_batch_run(vfps_batch* b, uint32_t nsteps);
/* the graphs vfps_batch_run(b, nsteps) replays, captured ahead */
VFPS_API int vfps_batch_prepare(vfps_batch* b, uint32_t nsteps);
VFPS_API int vfps_batch_record(vfps_batch* b, uint32_t capacity,
                               int want_profiles);
VFPS_API int vfps_batch_get_history(vfps_batch* b, uint32_t inst,
                                    uint32_t* count, float* csr_power,
                                    float* profiles);
VFPS_API int vfps_batch_get_mesh(vfps_batch* b, uint32_t inst, float* mesh,
                                 float* proj, float* wake);
/* the current x projection of every instance, concatenated (sum of n) */
VFPS_API int vfps_batch_get_projections(vfps_batch* b, float* out);
/* recorded CSR power of steps [first, first+count) of every instance,
 * [ninst][count] */
VFPS_API int vfps_batch_get_csr_histories(vfps_batch* b, uint32_t first,
                                          uint32_t count, float* out);
/* instances, launch groups and mesh cells per step */
VFPS_API int vfps_batch_info(vfps_batch* b, uint32_t* ninst,
                             uint32_t* ngroups, uint64_t* cells);
/* merge != 0: one launch per kernel variant, workgroup size and launch
 * form, whatever the tiles' LDS (each workgroup gets the largest); 0 (the
 * default): also one per LDS size class, the launches as parallel
 * branches.  Results do not depend on it. */
VFPS_API int vfps_batch_set_merge(vfps_batch* b, int merge);
/* vfps_batch_retile for every instance of mesh size n at once (one
 * rebuild of the launches, each in the form of a workgroup per tile or
 * persistent); on failure the previous tiling stays */
VFPS_API int vfps_batch_retile_size(vfps_batch* b, uint32_t n,
                                    uint32_t tile_rows, uint32_t threads,
                                    int persistent);
/* the launches of one step, cap rows of 4: workgroup size, tiles in the
 * block map (with padding), workgroups launched (fewer than the tiles:
 * persistent), dynamic LDS bytes; *ngroups = number of launches */
VFPS_API int vfps_batch_launches(vfps_batch* b, uint32_t* out, uint32_t cap,
                                 uint32_t* ngroups);
/* Set the tiling of one instance by hand: tile rows, workgroup size and
 * launch form (persistent != 0: as many workgroups as stay resident, each
 * walking several tiles; direct wake only).  Results do not depend on it.
 * On failure the previous tiling stays. */
VFPS_API int vfps_batch_retile(vfps_batch* b, uint32_t inst, uint32_t tile_rows,
                               uint32_t threads, int persistent);
/* The strip form for one instance (kick-drift maps with the direct wake):
 * a workgroup advances strip_tiles consecutive tiles of tile_rows rows
 * (at most 8 rows a tile, 64 a strip), kicking each row once and summing
 * the strip's x projection with one lane per row.  Results do not depend
 * on it.  On failure the previous tiling stays.  (No reference
 * counterpart: a launch shape of the batch.) */
VFPS_API int vfps_batch_retile_strip(vfps_batch* b, uint32_t inst,
                                     uint32_t tile_rows, uint32_t threads,
                                     uint32_t strip_tiles);
/* vfps_batch_retile_strip for every instance of mesh size n at once */
VFPS_API int vfps_batch_retile_size_strip(vfps_batch* b, uint32_t n,
                                          uint32_t tile_rows, uint32_t threads,
                                          uint32_t strip_tiles);
/* vfps_step_config of one instance (out[6], if nout >= 7: persistent;
 * out[7], if nout >= 8: tiles per strip, 0 outside the strip form) */
VFPS_API int vfps_batch_config(vfps_batch* b, uint32_t inst, uint32_t* out,
                               uint32_t nout);
/* average time of one batch step (all launches), HIP events on the device
 * stream; advances the state by nsteps */
VFPS_API int vfps_batch_timing(vfps_batch* b, uint32_t nsteps, float* ms_step);
/* Autotuner: per mesh size, time every tile height T in {1,2,4,8,16},
 * workgroup size in {128,256,512} and launch form (one workgroup per tile;
 * persistent: as many workgroups as stay resident, each walking several
 * tiles; for the kick-drift maps on meshes of 1024 rows or more also
 * strips of 16, 32 or 64 rows in tiles of 2, 4 or 8) that fits (steps steps each, on that size's
 * instances alone), time the three fastest again over 4 x steps and keep
 * the fastest of those.  The instances' state is restored afterwards;
 * results do not depend on the choice.  table (cap rows of VFPS_TUNE_COLS
 * floats: n, T, threads, us/step (of the longer timing where re-timed),
 * chosen, persistent, re-timed, tiles per strip) may be NULL. */
#define VFPS_TUNE_COLS 8
VFPS_API int vfps_batch_autotune(vfps_batch* b, uint32_t steps, float* table,
                                 uint32_t cap, uint32_t* nrows);

/* ---- learned surrogate (multilayer perceptron) --------------------------- *
 * No reference counterpart.  A fully connected network dims[0] -> dims[1]
 * -> ... -> dims[ndims-1] with GELU after every hidden layer, trained and
 * evaluated on the device: the products run on MFMA (BF16 operands, FP32
 * accumulation); master weights, gradients and Adam moments are FP32.
 * Outputs: dims[ndims-1]-1 regression values (mean squared error) and one
 * logit (binary cross entropy, weight label_weight).  Parameters are
 * flattened per layer as W [out][in] (row-major) then b [out]. */
VFPS_API int vfps_mlp_create(vfps_dev* dev, const uint32_t* dims,
                             uint32_t ndims, uint32_t max_batch,
                             uint64_t seed, vfps_mlp** out);
VFPS_API int vfps_mlp_destroy(vfps_mlp* m);
VFPS_API int vfps_mlp_nparams(const vfps_mlp* m, size_t* n);
/* set the FP32 parameters (their BF16 copy follows); resets Adam */
VFPS_API int vfps_mlp_set_params(vfps_mlp* m, const float* params, size_t n);
/* parameters and the gradients of the last vfps_mlp_loss_grad (either may
 * be NULL) */
VFPS_API int vfps_mlp_get_params(vfps_mlp* m, float* params, float* grads,
                                 size_t n);
/* y [B][dims_out] = network(x [B][dims_in]) */
VFPS_API int vfps_mlp_forward(vfps_mlp* m, const float* x, uint32_t B,
                              float* y);
/* loss of one batch and its gradients (kept on the device) */
VFPS_API int vfps_mlp_loss_grad(vfps_mlp* m, const float* x,
                                const float* target, const float* label,
                                uint32_t B, float label_weight, float* loss);
/* one Adam step with the current gradients */
VFPS_API int vfps_mlp_adam_step(vfps_mlp* m, float lr, float beta1,
                                float beta2, float eps);
/* epochs of shuffled minibatch Adam (beta1 0.9, beta2 0.999, eps 1e-8) on N
 * samples held on the device; epoch_loss (epochs) may be NULL */
VFPS_API int vfps_mlp_train(vfps_mlp* m, const float* x, const float* target,
                            const float* label, uint32_t N, uint32_t batch,
                            uint32_t epochs, float lr, float label_weight,
                            uint64_t seed, float* epoch_loss);
/* the layers' product alone: C [M][N] = A B, A(i,k) = A[i*asi + k*ask],
 * B(k,j) = B[k*bsk + j*bsj], operands rounded to BF16 */
VFPS_API int vfps_gemm_bf16(vfps_dev* dev, uint32_t M, uint32_t N, uint32_t K,
                            const float* A, size_t a_len, int64_t asi,
                            int64_t ask, const float* B, size_t b_len,
                            int64_t bsk, int64_t bsj, float* C);
/* the BF16 rounding of the kernels (round to nearest even), host only */
VFPS_API int vfps_round_bf16(const float* x, float* y, size_t n);

#ifdef __cplusplus
}
#endif

#endif
