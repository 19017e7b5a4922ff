/*
 * vfps_hdf5.h -- result and checkpoint files in the layout of Inovesa's
 * HDF5File (inc/IO/HDF5File.hpp, src/IO/HDF5File.cpp), host only.
 *
 * A separate library (inovesa_amd/lib/libvfps_h5.so, over the HDF5 C
 * library) so that libvfps.so does not depend on HDF5.  The caller hands
 * host arrays (downloaded from the device objects); every call returns 0 on
 * success and a negative vfps_status on failure, with vfps_h5_last_error().
 * All data sets hold 32-bit floats (IEEE_F32LE) as the reference's
 * meshaxis_t/integral_t are float; the bucket numbers are NATIVE_UINT32.
 */
#ifndef VFPS_HDF5_H
#define VFPS_HDF5_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFPS_H5_API __attribute__((visibility("default")))

typedef struct vfps_h5 vfps_h5;

/* What the constructor HDF5File(filename, ps, ef, imp, nparticles, t_sync,
 * f_rev) (src/IO/HDF5File.cpp:29-339) takes from its objects. */
typedef struct {
    uint32_t nb;              /* PhaseSpace::nb */
    uint32_t nx, ny;          /* PhaseSpace::nx, ny */
    uint32_t nbuckets;        /* ef->getBuckets().size(); 0: no field */
    const uint32_t* buckets;  /* the bucket numbers */
    uint32_t maxn;            /* ef->getNMax()/2: CSR spectrum bins */
    uint32_t imp_size;        /* imp->nFreqs()/2; 0: no impedance */
    uint32_t nparticles;      /* tracked particles */
    const float* axis_z;      /* ps->getAxis(0)->data(), nx values */
    const float* axis_f;      /* frequency ruler, maxn values (may be NULL
                                 when maxn == 0) */
    double meter;             /* ps->getScale(0,"Meter") */
    double electronvolt;      /* ps->getScale(1,"ElectronVolt") */
    double hertz;             /* the frequency ruler's "Hertz" scale */
    double t_sync, f_rev;
    double current, charge;   /* ps->current, ps->charge */
    double volts;             /* ef->volts */
    double factor4WattPerHertz, factor4Watts;
    double factor4Ohms;       /* Impedance::factor4Ohms */
    const float* imp_real;    /* impedance, imp_size values each */
    const float* imp_imag;
    int32_t version[3];       /* /Info/Inovesa_v */
    const char* build;        /* /Info/Inovesa_build (may be NULL) */
} vfps_h5_info;

VFPS_H5_API const char* vfps_h5_last_error(void);

/* HDF5File::HDF5File: truncates and lays out the file */
VFPS_H5_API int vfps_h5_create(const char* path, const vfps_h5_info* info,
                               vfps_h5** out);
VFPS_H5_API int vfps_h5_close(vfps_h5* f);

/* HDF5File::AppendType (inc/IO/HDF5File.hpp:98-100) */
enum { VFPS_H5_ALL = 0, VFPS_H5_DEFAULTS = 1, VFPS_H5_PHASESPACE = 2 };

/* append(const PhaseSpace&, t, at) (src/IO/HDF5File.cpp:391-419).
 * mesh [nb][nx][ny] is read for ALL and PHASESPACE; the rest ([nb][nx]
 * projections, [nb] moments) for ALL and DEFAULTS (may be NULL otherwise):
 * xproj = getProjection(0), yproj = getProjection(1), length =
 * getBunchLength(), position = getMoment(0,0), spread = getEnergySpread(),
 * average = getMoment(1,0), population = getBunchPopulation(). */
VFPS_H5_API int vfps_h5_append_ps(vfps_h5* f, float t, int at,
                                  const float* mesh, const float* xproj,
                                  const float* yproj, const float* length,
                                  const float* position, const float* spread,
                                  const float* average,
                                  const float* population);
/* append(const ElectricField*, fullspectrum) (:341-347): power [nb],
 * spectrum [nb][maxn] (NULL: fullspectrum false) */
VFPS_H5_API int vfps_h5_append_csr(vfps_h5* f, const float* power,
                                   const float* spectrum);
/* appendPadded (:349-353): imp_size values of each */
VFPS_H5_API int vfps_h5_append_padded(vfps_h5* f, const float* profile,
                                      const float* wake);
/* appendRFKicks (:355-359): count rows of (phase, amplitude) */
VFPS_H5_API int vfps_h5_append_rfkicks(vfps_h5* f, const float* kicks,
                                       uint32_t count);
/* appendTracks (:361-370): [nparticles][2] positions in physical units */
VFPS_H5_API int vfps_h5_append_tracks(vfps_h5* f, const float* positions);
/* append(const WakeKickMap*) (:421-424): getForce(), [nb][nx] */
VFPS_H5_API int vfps_h5_append_wake(vfps_h5* f, const float* wake);
/* addParameterToGroup (:332-339) with a double, an int64 or a string */
VFPS_H5_API int vfps_h5_add_parameter_f64(vfps_h5* f, const char* group,
                                          const char* name, double value);
VFPS_H5_API int vfps_h5_add_parameter_i64(vfps_h5* f, const char* group,
                                          const char* name, int64_t value);
VFPS_H5_API int vfps_h5_add_parameter_str(vfps_h5* f, const char* group,
                                          const char* name,
                                          const char* value);
/* addParameterToGroup with the data type ProgramOptions::save(HDF5File*)
 * (src/IO/ProgramOptions.cpp:460-490) gives each option: value points to a
 * float, double, uint32_t or int32_t */
enum { VFPS_H5_F32 = 0, VFPS_H5_F64 = 1, VFPS_H5_U32 = 2, VFPS_H5_I32 = 3 };
VFPS_H5_API int vfps_h5_add_parameter_typed(vfps_h5* f, const char* group,
                                            const char* name, int type,
                                            const void* value);
/* number of rows appended so far to the named data set (e.g.
 * "/PhaseSpace/data") */
VFPS_H5_API int vfps_h5_rows(vfps_h5* f, const char* dataset,
                             uint64_t* rows);

/* HDF5File::readPhaseSpace (src/IO/HDF5File.cpp:428-495): phase space
 * number `step` of /PhaseSpace/data (negative: from the end, -1 the last).
 * Rank 3 files ([t][x][y]) hold one bunch, rank 4 ([t][b][x][y]) nb.
 * out may be NULL to query nb and n; otherwise it receives [nb][n][n]. */
VFPS_H5_API int vfps_h5_read_phasespace(const char* path, int64_t step,
                                        uint32_t* nb, uint32_t* n,
                                        uint64_t* nsteps, float* out);

#ifdef __cplusplus
}
#endif

#endif
