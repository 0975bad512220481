/*
 * ql_io.h -- the reference's on-disk formats for the QL state, in a library of
 * its own (libql_io.so, HDF5 1.10) so that the GPU library does not depend on
 * HDF5.  All functions work on host buffers in the layouts of ql_dns.h and
 * return 0 or a negative code with qlio_last_error().
 *
 * FullSolution.h5 (Auxiliary/FullSave_Load.cpp:11-362):
 *   /ky                 [nxy] float   Im(ky) per mode (FullSave_Load.cpp:117-134)
 *   /kz                 [nz] double   Im(kz) per kz slot (:137-148)
 *   /t=<t as %.15g>/    one group per save (:183-188)
 *     "Ind jj: kx I ky J" [nz*nLin][2] float   mode jj (kx index I, ky index J,
 *                                              Kdata.cpp:28-46), fields
 *                                              (ux, zeta, bx, eta) one after
 *                                              another, {re, im} (:191-214)
 *     "Mean"              [nz*nMF][2]  float   mean fields (:219-225)
 *     "kx"                [nxy] float          Im(kx) + q t Im(ky) (:231-240)
 * The reference stores float32; with fp64 != 0 the same datasets are written
 * as float64 and the group also gets "kx0" [nxy] double (Im(kx) without the
 * shear term) and a double attribute "t", so that a restart from it
 * reproduces the state bit for bit.  A restart (FullSave_Load.cpp:264-347)
 * reads the group with the largest t, and kx0 = kx - q t ky unless "kx0" is
 * present.
 *
 * Time series (Auxiliary/TimeVariables.cpp:52-228): raw little-endian doubles
 * appended per save to energy.dat (4), angular_momentum.dat (4),
 * dissipation.dat (4), reynolds_stress.dat (nrey), time_vec.dat (1) and
 * mean_fields.dat (nMF x NZ real-space profiles).
 */
#ifndef QL_IO_H
#define QL_IO_H

#include "ql_dns.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qlio_layout {
    int NX, NY, NZ;       /* grid */
    int nxy, nz;          /* modes and dealiased kz slots (Models/Model.h:32-36) */
    int nlin, nmf;        /* fluctuation and mean fields */
    double L[3];          /* box */
} qlio_layout;

/* layout of a run with these parameters */
int qlio_layout_from_params(const ql_params *p, qlio_layout *out);

/* FullSave_Load::FullSave_Load + SaveSolution (FullSave_Load.cpp:11-252):
 * create != 0 truncates the file and writes /ky, /kz first (otherwise the
 * file is opened, or created if missing); then the group for time t. */
int qlio_h5_save(const char *path, int create, const qlio_layout *L, double t, double q, const double *lin,
                 const double *mf, const double *kx0, int fp64);

/* FullSave_Load::LoadSolution_forRestart (FullSave_Load.cpp:264-347): the
 * save with the largest t -> *t, lin, mf, kx0 (any output may be NULL). */
int qlio_h5_load_latest(const char *path, const qlio_layout *L, double q, double *t, double *lin, double *mf,
                        double *kx0);

/* number of saved time groups in the file */
int qlio_h5_count(const char *path, int *ngroups);

/* TimeVariables: truncate (or create) the six .dat files in dir
 * (TimeVariables.cpp:52-118) */
int qlio_dat_init(const char *dir);
/* TimeVariables::Save_Data + Save_Mean_Fields (TimeVariables.cpp:151-228):
 * diag as ql_diagnostics returns it (energy 0..3, angular momentum 4..7,
 * dissipation 8..11, Reynolds 12..15); nrey values of reynolds are written
 * (values past the 4 diag holds are written as 0); mf_real [nmf][NZ] may be
 * NULL to skip mean_fields.dat. */
int qlio_dat_append(const char *dir, double t, const double diag[18], int nrey, const double *mf_real, int nmf,
                    int NZ);

/* the same from TimeVariables' arrays: energy, angular momentum and
 * dissipation (4 each) and nrey Reynolds values (HD_fullU saves 2 NZ stress
 * profile values, HD_fullU.cpp:524-535) */
int qlio_dat_append_values(const char *dir, double t, const double energy[4], const double am[4],
                           const double diss[4], const double *reynolds, int nrey, const double *mf_real, int nmf,
                           int NZ);

const char *qlio_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
