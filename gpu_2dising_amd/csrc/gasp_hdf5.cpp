// gasp_hdf5.cpp -- the HDF5 grid files of the Python front-end
// (gridstates.hdf5 for nucleation swarms, pBgrids.hdf5 for committor runs).
//
// File layout (src/io.c:132-301, written by src/ising_ext.cu:360 at every
// grid output):
//   /L, /ngrids, /tot_nsweeps                int scalars
//   /h, /beta                                double scalars
//   /total_saved_grids                       int scalar, grids written so far
//   /sweep_<isweep>/grid_<g>                 int [L][L], one per replica
//     attributes (variable-length strings): magnetisation ("%g" or "null"),
//     lclus_size ("%g" or "null"), committor ("null"), committor_error ("null")
// A sweep group in which no grid could be written is removed again.
// Errors print to stderr and return -1, as the reference does.
#include <hdf5.h>
#include <stdio.h>

#include "../../include/gasp_hdf5.h"

namespace {

// Closes an HDF5 identifier with the matching H5?close on scope exit.
struct Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Handle() {
        if (id >= 0) close(id);
    }
    Handle(const Handle &) = delete;
    Handle &operator=(const Handle &) = delete;
    bool ok() const { return id >= 0; }
};

bool write_scalar(hid_t file, hid_t space, const char *name, hid_t type, const void *value, bool fatal) {
    Handle d(H5Dcreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!d.ok()) {
        fprintf(stderr, fatal ? "Error: failed to create dataset '%s'\n" : "Warning: failed to create dataset '%s'\n",
                name);
        return !fatal;
    }
    if (H5Dwrite(d.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, value) < 0) {
        fprintf(stderr, fatal ? "Error: failed to write dataset '%s'\n" : "Warning: failed to initialize '%s'\n",
                name);
        return !fatal;
    }
    return true;
}

void string_attr(hid_t obj, const char *name, hid_t str_type, hid_t scalar, const char *text) {
    Handle a(H5Acreate2(obj, name, str_type, scalar, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.ok()) H5Awrite(a.id, str_type, &text);
}

}  // namespace

extern "C" int create_ising_grids_hdf5(int L, int ngrids, int tot_nsweeps, double h, double beta, int itask,
                                       char *filename) {
    (void)itask;  // not recorded in the file
    Handle file(H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) {
        fprintf(stderr, "Error: failed to create HDF5 file '%s'\n", filename);
        return -1;
    }
    Handle scalar(H5Screate(H5S_SCALAR), H5Sclose);
    if (!scalar.ok()) {
        fprintf(stderr, "Error: failed to create HDF5 scalar dataspace\n");
        return -1;
    }
    const struct {
        const char *name;
        hid_t type;
        const void *value;
    } header[] = {{"L", H5T_NATIVE_INT, &L},
                  {"ngrids", H5T_NATIVE_INT, &ngrids},
                  {"tot_nsweeps", H5T_NATIVE_INT, &tot_nsweeps},
                  {"h", H5T_NATIVE_DOUBLE, &h},
                  {"beta", H5T_NATIVE_DOUBLE, &beta}};
    for (const auto &f : header)
        if (!write_scalar(file.id, scalar.id, f.name, f.type, f.value, true)) return -1;
    const int zero = 0;
    write_scalar(file.id, scalar.id, "total_saved_grids", H5T_NATIVE_INT, &zero, false);
    return 0;
}

extern "C" int write_ising_grids_hdf5(int L, int ngrids, int *ising_grids, int isweep, float *magnetisation,
                                      float *lclus_size, char *cv, double dn_thr, double up_thr, char *filename) {
    (void)cv;
    (void)dn_thr;
    (void)up_thr;  // every grid is written; the thresholds only filter the in-memory history
    Handle file(H5Fopen(filename, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) {
        fprintf(stderr, "Error: failed to open HDF5 file '%s' for writing\n", filename);
        return -1;
    }
    char group_name[64];
    snprintf(group_name, sizeof group_name, "/sweep_%d", isweep);
    int written = 0;
    {
        Handle group(H5Gcreate2(file.id, group_name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!group.ok()) {
            fprintf(stderr, "Error: failed to create group '%s' in HDF5 file\n", group_name);
            return -1;
        }
        Handle str_type(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(str_type.id, H5T_VARIABLE);
        Handle scalar(H5Screate(H5S_SCALAR), H5Sclose);
        const hsize_t dims[2] = {(hsize_t)L, (hsize_t)L};
        for (int g = 0; g < ngrids; ++g) {
            char name[64];
            snprintf(name, sizeof name, "grid_%d", g);
            Handle space(H5Screate_simple(2, dims, nullptr), H5Sclose);
            if (!space.ok()) {
                fprintf(stderr, "Error: failed to create dataspace for '%s'\n", name);
                continue;
            }
            Handle d(H5Dcreate2(group.id, name, H5T_NATIVE_INT, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                     H5Dclose);
            if (!d.ok()) {
                fprintf(stderr, "Error: failed to create dataset '%s'\n", name);
                continue;
            }
            if (H5Dwrite(d.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, ising_grids + (size_t)g * L * L) < 0)
                fprintf(stderr, "Error: failed to write dataset '%s'\n", name);
            else
                ++written;
            char mval[64], lval[64];
            const char *mtext = "null", *ltext = "null";
            if (magnetisation) {
                snprintf(mval, sizeof mval, "%g", (double)magnetisation[g]);
                mtext = mval;
            }
            if (lclus_size) {
                snprintf(lval, sizeof lval, "%g", (double)lclus_size[g]);
                ltext = lval;
            }
            string_attr(d.id, "magnetisation", str_type.id, scalar.id, mtext);
            string_attr(d.id, "lclus_size", str_type.id, scalar.id, ltext);
            string_attr(d.id, "committor", str_type.id, scalar.id, "null");
            string_attr(d.id, "committor_error", str_type.id, scalar.id, "null");
        }
    }
    if (written > 0) {
        Handle total(H5Dopen2(file.id, "total_saved_grids", H5P_DEFAULT), H5Dclose);
        int n = 0;
        H5Dread(total.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &n);
        n += written;
        H5Dwrite(total.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &n);
    } else {
        H5Ldelete(file.id, group_name, H5P_DEFAULT);
    }
    return 0;
}
