/* gasp_hdf5.h -- HDF5 grid files of the Python front-end (libgasp_hdf5.so,
 * linked against the HDF5 C library).  Same names, arguments, file layout and
 * error behaviour as the reference; only the Python module calls them (the
 * command line writes gridstates.bin). */
#ifndef GASP_HDF5_H
#define GASP_HDF5_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces create_ising_grids_hdf5 (include/io.h:14, src/io.c:132-193):
 * creates (truncates) `filename` with the scalar header datasets L, ngrids,
 * tot_nsweeps, h, beta and total_saved_grids = 0.  `itask` is not recorded.
 * Returns 0, or -1 after printing an error. */
int create_ising_grids_hdf5(int L, int ngrids, int tot_nsweeps, double h, double beta, int itask,
                            char *filename);

/* Replaces write_ising_grids_hdf5 (include/io.h:15, src/io.c:195-301): adds
 * group /sweep_<isweep> holding int [L][L] dataset grid_<g> for every replica,
 * with string attributes magnetisation / lclus_size ("%g", or "null" when the
 * array is NULL), committor and committor_error ("null"), and adds the number
 * of grids written to total_saved_grids (an empty group is deleted).  `cv` and
 * the thresholds are unused, as in the reference.  Returns 0, or -1 when the
 * file or group cannot be opened or created. */
int write_ising_grids_hdf5(int L, int ngrids, int *ising_grids, int isweep, float *magnetisation,
                           float *lclus_size, char *cv, double dn_thr, double up_thr, char *filename);

#ifdef __cplusplus
}
#endif

#endif /* GASP_HDF5_H */
