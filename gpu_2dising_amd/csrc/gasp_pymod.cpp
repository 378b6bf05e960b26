// gasp_pymod.cpp -- the PyGASP extension module `gasp` (src/ising_ext.cu) on
// top of the engine's C ABI.
//
// Same functions, keyword arguments, defaults, return values and module state
// as the reference:
//   run_nucleation_swarm(L, ngrid, tot_nsweeps, beta, h, initial_spin=-1,
//       cv='magnetisation', up_threshold=0.9, dn_threshold=-0.95,
//       mag_output_int=100, grid_output_int=1000, keep_grids=True,
//       max_keep_grids=0, threadsPerBlock=32, gpu_method=<auto>)
//       -> numpy.float32[tot_nsweeps // mag_output_int]
//   run_committor_calc(..., dn_threshold=-0.93, ..., grid_input='gridstates.bin',
//       grid_array=None, nsms=gpu_nsms) -> [(pB, bootstrap_error), ...]
//   grids, magnetisation, largest_cluster, gpu_nsms, GridSnapObject
// Differences (deliberate): snapshot arrays own their memory (the reference's
// alias a buffer freed by the next run); an explicit gpu_method is honoured
// (the reference always auto-selects); the default grid_input reads
// gridinput.bin instead of dereferencing a missing list; the GIL is released
// while the GPU runs and re-taken for output; keyword `seed` (default: the
// clock, as the reference) and `rng` ('philox' or 'mt19937') are added;
// without a GPU (import prints "Falling back to CPU", as the reference) the
// run functions use the reference's CPU engine, mc_driver_cpu with one MT19937
// stream seeded with `seed`, and `rng` / `gpu_method` do not apply;
// keyword `write_hdf5` (default True, as the reference always writes) can skip
// the HDF5 grid file (gridstates.hdf5 / pBgrids.hdf5, written through
// libgasp_hdf5 in the reference's layout).
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <structmember.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/gasp_amd.h"
#include "../../include/gasp_hdf5.h"

extern "C" void gasp_set_output(void (*cb)(const char *text), int style);

// ----------------------------------------------------------- module state
static int g_idev = -1;      // device chosen at import, -1 = none
static int g_gpu_nsms = 0;   // CU count (src/ising_ext.cu:55)
static PyObject *g_module = nullptr;

// =================================================================== GridSnapObject
// src/ising_ext.cu:74-207
typedef struct {
    PyObject_HEAD int isweep;
    int igrid;
    float magnetisation;
    float lclus_size;
    float committor;
    PyObject *grid;
} GridSnapObject;

static PyMemberDef GridSnap_members[] = {
    {"isweep", T_INT, offsetof(GridSnapObject, isweep), 0, "MC sweep at which the grid was captured"},
    {"igrid", T_INT, offsetof(GridSnapObject, igrid), 0, "Which grid in the ensemble is this"},
    {"magnetisation", T_FLOAT, offsetof(GridSnapObject, magnetisation), 0, "Magnetisation"},
    {"lclus_size", T_FLOAT, offsetof(GridSnapObject, lclus_size), 0,
     "Size of largest cluster of spins opposite to initial spin"},
    {"committor", T_FLOAT, offsetof(GridSnapObject, committor), 0, "Committor value if available. -1 if not."},
    {nullptr, 0, 0, 0, nullptr}};

static PyObject *GridSnap_get_grid(GridSnapObject *self, void *) {
    if (!self->grid) Py_RETURN_NONE;
    Py_INCREF(self->grid);
    return self->grid;
}

static int GridSnap_set_grid(GridSnapObject *self, PyObject *value, void *) {
    if (!value || !PyArray_Check(value)) {
        PyErr_SetString(PyExc_TypeError, "grid must be a numpy array");
        return -1;
    }
    Py_INCREF(value);
    Py_XSETREF(self->grid, value);
    return 0;
}

static PyGetSetDef GridSnap_getset[] = {
    {"grid", (getter)GridSnap_get_grid, (setter)GridSnap_set_grid, "2D numpy array of 8-bit integers", nullptr},
    {nullptr, nullptr, nullptr, nullptr, nullptr}};

static int GridSnap_init(GridSnapObject *self, PyObject *args, PyObject *) {
    int isweep = 0, igrid = 0;
    PyObject *arr = nullptr;
    if (!PyArg_ParseTuple(args, "iiO", &isweep, &igrid, &arr)) return -1;
    if (!PyArray_Check(arr)) {
        PyErr_SetString(PyExc_TypeError, "Third argument must be a NumPy array.");
        return -1;
    }
    PyArrayObject *a = (PyArrayObject *)arr;
    if (PyArray_TYPE(a) != NPY_INT8 || PyArray_NDIM(a) != 2) {
        PyErr_SetString(PyExc_TypeError, "NumPy array must be 2D with 8-bit integer elements.");
        return -1;
    }
    self->isweep = isweep;
    self->igrid = igrid;
    self->magnetisation = 0.0f;
    self->lclus_size = -1.0f;
    self->committor = -1.0f;
    Py_INCREF(arr);
    Py_XSETREF(self->grid, arr);
    return 0;
}

static void GridSnap_dealloc(GridSnapObject *self) {
    Py_XDECREF(self->grid);
    Py_TYPE(self)->tp_free((PyObject *)self);
}

static PyObject *GridSnap_deepcopy(PyObject *self, PyObject *memo);

static PyMethodDef GridSnap_methods[] = {
    {"__deepcopy__", (PyCFunction)GridSnap_deepcopy, METH_VARARGS, "Deep copy a GridSnapObject."},
    {nullptr, nullptr, 0, nullptr}};

static PyTypeObject GridSnapType = {PyVarObject_HEAD_INIT(nullptr, 0)};

static PyObject *GridSnap_deepcopy(PyObject *self, PyObject *) {
    GridSnapObject *o = (GridSnapObject *)self;
    PyObject *copy = PyObject_CallMethod(o->grid, "copy", nullptr);
    if (!copy) return nullptr;
    PyObject *args = Py_BuildValue("iiO", o->isweep, o->igrid, copy);
    Py_DECREF(copy);
    if (!args) return nullptr;
    PyObject *n = PyObject_CallObject((PyObject *)&GridSnapType, args);
    Py_DECREF(args);
    if (!n) return nullptr;
    GridSnapObject *c = (GridSnapObject *)n;
    c->magnetisation = o->magnetisation;
    c->lclus_size = o->lclus_size;
    c->committor = o->committor;
    return n;
}

// =================================================================== run state
struct RunState {
    bool write_hdf5 = true;
    bool keep_grids = false;
    long long maxhist = 0;
    long long ihist = 0;
    bool failed = false;  // a Python error was raised inside a callback
    PyObject *err_type = nullptr, *err_value = nullptr, *err_tb = nullptr;
};
static RunState g_run;

static void py_print(const char *text) {
    PyGILState_STATE st = PyGILState_Ensure();
    // PySys_WriteStdout truncates at 1000 bytes; driver lines are short
    PySys_WriteStdout("%s", text);
    PyGILState_Release(st);
}

static int replace_list(const char *name) {
    PyObject *l = PyList_New(0);
    if (!l) return -1;
    if (PyObject_SetAttrString(g_module, name, l) < 0) {
        Py_DECREF(l);
        return -1;
    }
    Py_DECREF(l);
    return 0;
}

// src/ising_ext.cu:216-275
static int reset_grids_list() {
    if (replace_list("grids") < 0 || replace_list("magnetisation") < 0 || replace_list("largest_cluster") < 0)
        return -1;
    return 0;
}

static PyObject *float_list(const float *v, int n) {
    PyObject *l = PyList_New(n);
    if (!l) return nullptr;
    for (int i = 0; i < n; ++i) {
        PyObject *x = PyFloat_FromDouble((double)v[i]);
        if (!x) {
            Py_DECREF(l);
            return nullptr;
        }
        PyList_SET_ITEM(l, i, x);
    }
    return l;
}

static int append_to_module_list(const char *name, PyObject *item) {
    PyObject *l = PyObject_GetAttrString(g_module, name);
    if (!l) return -1;
    int rc = PyList_Check(l) ? PyList_Append(l, item) : 0;
    Py_DECREF(l);
    return rc;
}

// Output callback (src/ising_ext.cu:280-463): module lists of magnetisation and
// largest cluster, every grid to the HDF5 file, and GridSnapObjects for
// replicas with dn_thr < cv < up_thr.
static int append_grids_list_body(int L, int ngrids, int *grid_data, int isweep, float *magnetisation,
                                  float *lclus_size, char *cv, double dn_thr, double up_thr, char *filename) {
    const float *cv_val = nullptr;
    if (strcmp(cv, "magnetisation") == 0) cv_val = magnetisation;
    else if (strcmp(cv, "largest_cluster") == 0) cv_val = lclus_size;
    if (!cv_val) {
        PyErr_SetString(PyExc_ValueError, "Invalid collective variable specified");
        return -1;
    }
    PyObject *ml = float_list(magnetisation, ngrids);
    if (!ml) return -1;
    int rc = append_to_module_list("magnetisation", ml);
    Py_DECREF(ml);
    if (rc < 0) return -1;
    if (lclus_size) {
        PyObject *cl = float_list(lclus_size, ngrids);
        if (!cl) return -1;
        rc = append_to_module_list("largest_cluster", cl);
        Py_DECREF(cl);
        if (rc < 0) return -1;
    }
    if (g_run.write_hdf5 &&
        write_ising_grids_hdf5(L, ngrids, grid_data, isweep, magnetisation, lclus_size, cv, dn_thr, up_thr,
                               filename) != 0) {
        PyErr_SetString(PyExc_RuntimeError, "Error writing HDF5 file");
        return -1;
    }
    if (!g_run.keep_grids || g_run.ihist >= g_run.maxhist) return 0;

    PyObject *snaps = PyList_New(0);
    if (!snaps) return -1;
    const int N = L * L;
    npy_intp dims[2] = {L, L};
    for (int i = 0; i < ngrids; ++i) {
        const double v = (double)cv_val[i];
        if (!(v > dn_thr && v < up_thr)) continue;
        PyObject *arr = PyArray_SimpleNew(2, dims, NPY_INT8);
        if (!arr) {
            Py_DECREF(snaps);
            return -1;
        }
        int8_t *dst = (int8_t *)PyArray_DATA((PyArrayObject *)arr);
        const int *src = grid_data + (size_t)i * N;
        for (int s = 0; s < N; ++s) dst[s] = (int8_t)src[s];
        PyObject *args = Py_BuildValue("iiO", isweep, i, arr);
        Py_DECREF(arr);
        if (!args) {
            Py_DECREF(snaps);
            return -1;
        }
        PyObject *snap = PyObject_CallObject((PyObject *)&GridSnapType, args);
        Py_DECREF(args);
        if (!snap) {
            Py_DECREF(snaps);
            return -1;
        }
        GridSnapObject *so = (GridSnapObject *)snap;
        so->magnetisation = magnetisation[i];
        if (lclus_size) so->lclus_size = lclus_size[i];
        so->committor = -1.0f;
        rc = PyList_Append(snaps, snap);
        Py_DECREF(snap);
        if (rc < 0) {
            Py_DECREF(snaps);
            return -1;
        }
        g_run.ihist++;
        if (g_run.ihist == g_run.maxhist) {
            PySys_WriteStderr("\nWARNING: Maximum history reached, not storing more snapshots to grids object!\n\n");
            g_run.ihist++;
            break;
        }
    }
    rc = 0;
    if (PyList_Size(snaps) > 0) rc = append_to_module_list("grids", snaps);
    Py_DECREF(snaps);
    return rc;
}

extern "C" int append_grids_list(int L, int ngrids, int *grid_data, int isweep, float *magnetisation,
                                 float *lclus_size, char *cv, double dn_thr, double up_thr, char *filename) {
    PyGILState_STATE st = PyGILState_Ensure();
    int rc = 0;
    if (!g_run.failed) {
        rc = append_grids_list_body(L, ngrids, grid_data, isweep, magnetisation, lclus_size, cv, dn_thr, up_thr,
                                    filename);
        if (rc < 0) {
            g_run.failed = true;
            PyErr_Fetch(&g_run.err_type, &g_run.err_value, &g_run.err_tb);
        }
    }
    PyGILState_Release(st);
    return rc;
}

// =================================================================== helpers
static bool parse_seed(PyObject *seed_obj, unsigned long *seed) {
    if (!seed_obj || seed_obj == Py_None) {
        *seed = (unsigned long)time(nullptr);
        return true;
    }
    unsigned long long v = PyLong_AsUnsignedLongLong(seed_obj);
    if (PyErr_Occurred()) return false;
    *seed = (unsigned long)v;
    return true;
}

static bool parse_method(PyObject *obj, int L, int ngrids, int tpb, int *method) {
    if (g_idev < 0) {  // the CPU engine has no methods
        *method = 0;
        return true;
    }
    if (!obj || obj == Py_None) {
        *method = select_gpu_method(L, ngrids, tpb, g_idev);
        return true;
    }
    long v = PyLong_AsLong(obj);
    if (PyErr_Occurred()) return false;
    *method = (int)v;
    return true;
}

static PyObject *raise_lib(const char *what) {
    PyErr_Format(PyExc_RuntimeError, "gasp: %s failed (error %d: %s)", what, gasp_last_error(), gasp_error_string());
    return nullptr;
}

// Runs the driver with the GIL released; returns false with a Python error set.
// Without a GPU (gpuInitDevice returned -1 at import) the reference's CPU
// engine runs instead: one MT19937 stream seeded with `seed`
// (src/ising_ext.cu:637-664, 1007-1030); `rng`, `tpb` and `method` do not apply.
static bool run_driver_cpu(int L, int ngrids, int *ising_grids, int *grid_fate, double beta, double h,
                           int tot_nsweeps, int mag_output_int, int grid_output_int, int itask, int initial_spin,
                           const char *cv, double dn, double up, int ninputs, float *result, unsigned long seed,
                           const char *filename) {
    std::string cvs(cv), fname(filename);
    mc_grids_t grids{L, ngrids, ising_grids};
    mc_sampler_t samples{tot_nsweeps, mag_output_int, grid_output_int};
    mc_function_t calc;
    calc.itask = itask;
    calc.dn_thr = dn;
    calc.up_thr = up;
    calc.ninputs = ninputs;
    calc.initial_spin = initial_spin;
    calc.cv = &cvs[0];
    calc.result = result;
    calc.filename = &fname[0];
    g_run.failed = false;
    g_run.err_type = g_run.err_value = g_run.err_tb = nullptr;
    init_genrand(seed);
    preComputeProbs_cpu(beta, h);
    if (g_run.write_hdf5) create_ising_grids_hdf5(L, ngrids, tot_nsweeps, h, beta, itask, &fname[0]);
    gasp_set_output(py_print, 1);
    int rc;
    Py_BEGIN_ALLOW_THREADS
    mc_driver_cpu(grids, beta, h, grid_fate, samples, calc, append_grids_list);
    rc = gasp_last_error();
    Py_END_ALLOW_THREADS
    gasp_set_output(nullptr, 0);
    if (g_run.failed) {
        PyErr_Restore(g_run.err_type, g_run.err_value, g_run.err_tb);
        g_run.err_type = g_run.err_value = g_run.err_tb = nullptr;
        return false;
    }
    if (rc) {
        raise_lib("mc_driver_cpu");
        return false;
    }
    return true;
}

static bool run_driver(int L, int ngrids, int *ising_grids, int *grid_fate, double beta, double h, int tot_nsweeps,
                       int mag_output_int, int grid_output_int, int itask, int initial_spin, const char *cv,
                       double dn, double up, int ninputs, float *result, int tpb, int method, unsigned long seed,
                       const char *rng, const char *filename) {
    if (g_idev < 0)
        return run_driver_cpu(L, ngrids, ising_grids, grid_fate, beta, h, tot_nsweeps, mag_output_int,
                              grid_output_int, itask, initial_spin, cv, dn, up, ninputs, result, seed, filename);
    int *d_grids = nullptr, *d_nb = nullptr;
    gpuInitGrid(L, ngrids, tpb, ising_grids, &d_grids, &d_nb);
    if (gasp_last_error()) {
        raise_lib("gpuInitGrid");
        return false;
    }
    gasp_rng_state *d_state = nullptr;
    if (strcmp(rng, "mt19937") == 0) {
        gasp_init_rand_mt(ngrids, seed, &d_state);
    } else if (strcmp(rng, "philox") == 0) {
        gpuInitRand(ngrids, tpb, seed, &d_state);
    } else {
        gasp_free(d_grids);
        gasp_free(d_nb);
        PyErr_SetString(PyExc_ValueError, "rng must be 'philox' or 'mt19937'");
        return false;
    }
    if (gasp_last_error()) {
        gasp_free(d_grids);
        gasp_free(d_nb);
        raise_lib("gpuInitRand");
        return false;
    }
    preComputeProbs_gpu(beta, h);
    std::string cvs(cv), fname(filename);
    mc_gpu_grids_t grids{L, ngrids, ising_grids, d_grids, d_nb};
    mc_sampler_t samples{tot_nsweeps, mag_output_int, grid_output_int};
    mc_function_t calc;
    calc.itask = itask;
    calc.dn_thr = dn;
    calc.up_thr = up;
    calc.ninputs = ninputs;
    calc.initial_spin = initial_spin;
    calc.cv = &cvs[0];
    calc.result = result;
    calc.filename = &fname[0];
    gpu_run_t run{d_state, tpb, method};

    g_run.failed = false;
    g_run.err_type = g_run.err_value = g_run.err_tb = nullptr;
    // src/ising_ext.cu:660, 700, 1026, 1062 (return value unchecked there too:
    // a file that cannot be created fails at the first grid output)
    if (g_run.write_hdf5) create_ising_grids_hdf5(L, ngrids, tot_nsweeps, h, beta, itask, &fname[0]);
    gasp_set_output(py_print, 1);
    int rc;
    Py_BEGIN_ALLOW_THREADS
    mc_driver_gpu(grids, beta, h, grid_fate, samples, calc, run, append_grids_list);
    rc = gasp_last_error();
    Py_END_ALLOW_THREADS
    gasp_set_output(nullptr, 0);
    gasp_free(d_state);
    gasp_free(d_grids);
    gasp_free(d_nb);
    if (g_run.failed) {
        PyErr_Restore(g_run.err_type, g_run.err_value, g_run.err_tb);
        g_run.err_type = g_run.err_value = g_run.err_tb = nullptr;
        return false;
    }
    if (rc) {
        raise_lib("mc_driver_gpu");
        return false;
    }
    return true;
}

static bool begin_history(int L, int ngrids, int tot_nsweeps, int grid_output_int, int keep_grids,
                          int max_keep_grids) {
    // src/ising_ext.cu:587-613
    remove("gridstates.bin");
    if (reset_grids_list() < 0) return false;
    const long long nsnaps = grid_output_int > 0 ? tot_nsweeps / grid_output_int + 1 : 1;
    g_run.keep_grids = keep_grids != 0;
    g_run.maxhist = max_keep_grids ? max_keep_grids : nsnaps * ngrids;
    g_run.ihist = 0;
    (void)L;
    return true;
}

// =================================================================== run_nucleation_swarm
static const char *nuc_doc =
    "run_nucleation_swarm(L, ngrid, tot_nsweeps, beta, h, initial_spin=-1, cv='magnetisation', up_threshold=0.9,\n"
    "                     dn_threshold=-0.95, mag_output_int=100, grid_output_int=1000, keep_grids=True,\n"
    "                     max_keep_grids=0, threadsPerBlock=32, gpu_method=None, seed=None, rng='philox',\n"
    "                     write_hdf5=True)\n"
    "\n"
    "Simulate a swarm of 2D Ising replicas from a uniform initial_spin state on the GPU.  Returns the\n"
    "fraction of replicas whose collective variable exceeds up_threshold at each mag_output_int\n"
    "(numpy.float32, length tot_nsweeps // mag_output_int; 1.0 after all have nucleated).  Every\n"
    "grid_output_int sweeps the module lists `magnetisation` (and `largest_cluster`) get one entry\n"
    "per replica, and `grids` gets the GridSnapObjects of replicas with dn_threshold < cv < up_threshold\n"
    "(at most max_keep_grids in total; 0 means (tot_nsweeps // grid_output_int + 1) * ngrid).\n"
    "gpu_method=None selects the layout with select_gpu_method; rng='mt19937' uses per-replica\n"
    "MT19937 streams seeded seed + r (the reference CPU stream).";

static PyObject *run_nucleation_swarm(PyObject *, PyObject *args, PyObject *kwargs) {
    int L = 64, ngrids = 128, tot_nsweeps = 100;
    double beta = 0.54, h = 0.07;
    int initial_spin = -1;
    // defaults fixed before parsing, as src/ising_ext.cu:537-540
    double up_threshold = -0.90 * (double)initial_spin;
    double dn_threshold = 0.95 * (double)initial_spin;
    int mag_output_int = 100, grid_output_int = 1000, keep_grids = 1, max_keep_grids = 0, tpb = 32;
    PyObject *method_obj = Py_None, *seed_obj = Py_None;
    const char *cv = "magnetisation";
    const char *rng = "philox";
    static const char *kwlist[] = {"L", "ngrid", "tot_nsweeps", "beta", "h", "initial_spin", "cv", "up_threshold",
                                   "dn_threshold", "mag_output_int", "grid_output_int", "keep_grids",
                                   "max_keep_grids", "threadsPerBlock", "gpu_method", "seed", "rng",
                                   "write_hdf5", nullptr};
    int write_hdf5 = 1;
    if (!PyArg_ParseTupleAndKeywords(args, kwargs, "iiidd|isddiipiiOOsp", (char **)kwlist, &L, &ngrids,
                                     &tot_nsweeps, &beta, &h, &initial_spin, &cv, &up_threshold, &dn_threshold,
                                     &mag_output_int, &grid_output_int, &keep_grids, &max_keep_grids, &tpb,
                                     &method_obj, &seed_obj, &rng, &write_hdf5))
        return nullptr;
    g_run.write_hdf5 = write_hdf5 != 0;
    if (strcmp(cv, "magnetisation") != 0 && strcmp(cv, "largest_cluster") != 0) {
        PyErr_SetString(PyExc_ValueError, "cv must be either 'magnetisation' or 'largest_cluster'");
        return nullptr;
    }
    if (strcmp(cv, "largest_cluster") == 0 && (fabs(up_threshold) < 2.0 || fabs(dn_threshold) < 2.0)) {
        if (PyErr_WarnEx(PyExc_UserWarning,
                         "For cv='largest_cluster', up_threshold and dn_threshold should have magnitudes >= 2", 1) < 0)
            return nullptr;
    }
    if (max_keep_grids < 0) {
        PyErr_SetString(PyExc_ValueError, "max_keep_grids must be a positive integer");
        return nullptr;
    }
    if (L < 2 || ngrids < 1 || tot_nsweeps < 0 || mag_output_int < 1 || grid_output_int < 1) {
        PyErr_SetString(PyExc_ValueError, "L >= 2, ngrid >= 1, mag_output_int >= 1 and grid_output_int >= 1 required");
        return nullptr;
    }
    unsigned long seed;
    if (!parse_seed(seed_obj, &seed)) return nullptr;
    int method;
    if (!parse_method(method_obj, L, ngrids, tpb, &method)) return nullptr;
    if (!begin_history(L, ngrids, tot_nsweeps, grid_output_int, keep_grids, max_keep_grids)) return nullptr;

    int *ising_grids = init_grids_uniform(L, ngrids, initial_spin);
    if (!ising_grids) return PyErr_NoMemory();
    const int nres = tot_nsweeps / mag_output_int;
    std::vector<float> result(nres > 0 ? nres : 1, 0.0f);
    const bool ok = run_driver(L, ngrids, ising_grids, nullptr, beta, h, tot_nsweeps, mag_output_int, grid_output_int,
                               0, initial_spin, cv, dn_threshold, up_threshold, 1, result.data(), tpb, method, seed,
                               rng, "gridstates.hdf5");
    free(ising_grids);
    if (!ok) return nullptr;
    npy_intp dims[1] = {nres};
    PyObject *out = PyArray_SimpleNew(1, dims, NPY_FLOAT32);
    if (!out) return nullptr;
    if (nres > 0) memcpy(PyArray_DATA((PyArrayObject *)out), result.data(), (size_t)nres * sizeof(float));
    return out;
}

// =================================================================== run_committor_calc
static const char *comm_doc =
    "run_committor_calc(L, ngrid, tot_nsweeps, beta, h, initial_spin=-1, cv='magnetisation', up_threshold=0.9,\n"
    "                   dn_threshold=-0.93, mag_output_int=100, grid_output_int=1000, keep_grids=True,\n"
    "                   max_keep_grids=0, threadsPerBlock=32, gpu_method=None, grid_input='gridstates.bin',\n"
    "                   grid_array=None, nsms=gpu_nsms, seed=None, rng='philox', write_hdf5=True,\n"
    "                   bootstrap='libc')\n"
    "\n"
    "Committor by shooting: each input grid (grid_input='NumPy' with grid_array a list of LxL int8\n"
    "arrays, or gridinput.bin) is copied into ngrid // len(grid_array) replicas, which run until their\n"
    "collective variable exceeds up_threshold (fate B) or falls below dn_threshold (fate A), or\n"
    "tot_nsweeps.  Returns [(pB, bootstrap standard error), ...], one tuple per input grid.\n"
    "bootstrap='libc' draws the resamples from libc rand() as the reference does; 'gpu' runs the\n"
    "same estimator on the GPU with Philox draws (statistically equal, much faster).";

static PyObject *run_committor_calc(PyObject *, PyObject *args, PyObject *kwargs) {
    int L = 64, ngrids = 128, tot_nsweeps = 100;
    double beta = 0.54, h = 0.07;
    int initial_spin = -1;
    double up_threshold = -0.90 * (double)initial_spin;  // src/ising_ext.cu:821-822
    double dn_threshold = 0.93 * (double)initial_spin;
    int mag_output_int = 100, grid_output_int = 1000, keep_grids = 1, max_keep_grids = 0, tpb = 32;
    PyObject *method_obj = Py_None, *seed_obj = Py_None, *grid_array_obj = nullptr;
    const char *cv = "magnetisation";
    const char *grid_input = "gridstates.bin";
    const char *rng = "philox";
    int nsms = g_gpu_nsms;
    static const char *kwlist[] = {"L", "ngrid", "tot_nsweeps", "beta", "h", "initial_spin", "cv", "up_threshold",
                                   "dn_threshold", "mag_output_int", "grid_output_int", "keep_grids",
                                   "max_keep_grids", "threadsPerBlock", "gpu_method", "grid_input", "grid_array",
                                   "nsms", "seed", "rng", "write_hdf5", "bootstrap", nullptr};
    int write_hdf5 = 1;
    const char *bootstrap = "libc";
    if (!PyArg_ParseTupleAndKeywords(args, kwargs, "iiidd|isddiipiiOsOiOsps", (char **)kwlist, &L, &ngrids,
                                     &tot_nsweeps, &beta, &h, &initial_spin, &cv, &up_threshold, &dn_threshold,
                                     &mag_output_int, &grid_output_int, &keep_grids, &max_keep_grids, &tpb,
                                     &method_obj, &grid_input, &grid_array_obj, &nsms, &seed_obj, &rng,
                                     &write_hdf5, &bootstrap))
        return nullptr;
    if (strcmp(bootstrap, "libc") != 0 && strcmp(bootstrap, "gpu") != 0) {
        PyErr_SetString(PyExc_ValueError, "bootstrap must be 'libc' or 'gpu'");
        return nullptr;
    }
    g_run.write_hdf5 = write_hdf5 != 0;
    // src/ising_ext.cu:852-900
    if (nsms < 0) {
        PyErr_SetString(PyExc_ValueError, "nsms must be a non-negative integer");
        return nullptr;
    }
    g_gpu_nsms = nsms;
    if (strcmp(cv, "magnetisation") != 0 && strcmp(cv, "largest_cluster") != 0) {
        PyErr_SetString(PyExc_ValueError, "cv must be either 'magnetisation' or 'largest_cluster'");
        return nullptr;
    }
    if (max_keep_grids < 0) {
        PyErr_SetString(PyExc_ValueError, "max_keep_grids must be a non-negative integer");
        return nullptr;
    }
    if (L < 2 || ngrids < 1 || tot_nsweeps < 0 || mag_output_int < 1 || grid_output_int < 1) {
        PyErr_SetString(PyExc_ValueError, "L >= 2, ngrid >= 1, mag_output_int >= 1 and grid_output_int >= 1 required");
        return nullptr;
    }
    const bool have_array = grid_array_obj && grid_array_obj != Py_None;
    int count = 1;
    if (have_array) {
        if (!PyList_Check(grid_array_obj)) {
            PyErr_SetString(PyExc_TypeError, "grid_array must be a list of 2D NumPy arrays");
            return nullptr;
        }
        count = (int)PyList_Size(grid_array_obj);
        if (count > g_gpu_nsms) {
            PyErr_SetString(PyExc_ValueError, "grid_array list length exceeds gpu_nsms");
            return nullptr;
        }
        if (count == 0) {
            PyErr_SetString(PyExc_ValueError, "grid_array list is empty");
            return nullptr;
        }
        if (g_gpu_nsms % count != 0) {
            PyErr_SetString(PyExc_ValueError, "grid_array list length must divide gpu_nsms exactly");
            return nullptr;
        }
        if (ngrids % g_gpu_nsms != 0) {
            PyErr_SetString(PyExc_ValueError, "ngrids must be a multiple of gpu_nsms");
            return nullptr;
        }
        if (ngrids % count != 0) {
            PyErr_SetString(PyExc_ValueError, "ngrids must be a multiple of grid_array list length");
            return nullptr;
        }
    }
    std::vector<int> grid_array_c;
    if (have_array) {
        grid_array_c.resize((size_t)count * L * L);
        for (int k = 0; k < count; ++k) {
            PyObject *item = PyList_GetItem(grid_array_obj, k);
            if (!PyArray_Check(item)) {
                PyErr_SetString(PyExc_TypeError, "All items in grid_array must be NumPy arrays");
                return nullptr;
            }
            PyArrayObject *a = (PyArrayObject *)item;
            if (PyArray_TYPE(a) != NPY_INT8 || PyArray_NDIM(a) != 2) {
                PyErr_SetString(PyExc_TypeError, "Each grid_array item must be a 2D NumPy array of type int8");
                return nullptr;
            }
            if (PyArray_DIM(a, 0) != L || PyArray_DIM(a, 1) != L) {
                PyErr_SetString(PyExc_ValueError, "Each grid_array item must have shape LxL");
                return nullptr;
            }
            PyArrayObject *c = (PyArrayObject *)PyArray_GETCONTIGUOUS(a);
            if (!c) return nullptr;
            const int8_t *src = (const int8_t *)PyArray_DATA(c);
            for (int s = 0; s < L * L; ++s) grid_array_c[(size_t)k * L * L + s] = (int)src[s];
            Py_DECREF(c);
        }
    }
    unsigned long seed;
    if (!parse_seed(seed_obj, &seed)) return nullptr;
    int method;
    if (!parse_method(method_obj, L, ngrids, tpb, &method)) return nullptr;

    int *ising_grids = nullptr;
    if (strcmp(grid_input, "gridstates.bin") == 0 && !have_array) {
        ising_grids = init_grids_from_file(L, ngrids);  // reads gridinput.bin (src/io.c:11)
        if (!ising_grids) return raise_lib("init_grids_from_file");
    } else if (strcmp(grid_input, "NumPy") == 0 && have_array) {
        ising_grids = init_grids_from_array(L, ngrids, count, grid_array_c.data());
        if (!ising_grids) return raise_lib("init_grids_from_array");
    } else {
        PyErr_SetString(PyExc_ValueError, "Invalid grid_input option or no grid_array supplied");
        return nullptr;
    }
    if (!begin_history(L, ngrids, tot_nsweeps, grid_output_int, keep_grids, max_keep_grids)) {
        free(ising_grids);
        return nullptr;
    }
    int *fates = init_fates(ngrids);
    if (!fates) {
        free(ising_grids);
        return PyErr_NoMemory();
    }
    std::vector<float> result(count, 0.0f);
    const bool ok = run_driver(L, ngrids, ising_grids, fates, beta, h, tot_nsweeps, mag_output_int, grid_output_int,
                               1, initial_spin, cv, dn_threshold, up_threshold, count, result.data(), tpb, method,
                               seed, rng, "pBgrids.hdf5");
    free(ising_grids);
    if (!ok) {
        free(fates);
        return nullptr;
    }
    // src/ising_ext.cu:1077-1088
    const int per = ngrids / count;
    std::vector<double> err(count, 0.0);
    if (strcmp(bootstrap, "gpu") == 0) {
        // same estimator on the GPU, Philox draws instead of libc rand()
        if (gasp_bootstrap_errbar_gpu(fates, 0, count, per, per, (unsigned long long)seed, err.data()) != 0) {
            free(fates);
            raise_lib("gasp_bootstrap_errbar_gpu");
            return nullptr;
        }
    } else {
        for (int i = 0; i < count; ++i) {
            err[i] = bootstrap_errbar(per, per, fates + (size_t)i * per);
            if (err[i] < 0.0) {
                free(fates);
                PyErr_SetString(PyExc_RuntimeError, "Error calculating bootstrap error bar");
                return nullptr;
            }
        }
    }
    free(fates);
    PyObject *out = PyList_New(count);
    if (!out) return nullptr;
    for (int i = 0; i < count; ++i) {
        PyObject *t = Py_BuildValue("(dd)", (double)result[i], err[i]);
        if (!t) {
            Py_DECREF(out);
            return nullptr;
        }
        PyList_SET_ITEM(out, i, t);
    }
    return out;
}

// =================================================================== module
static PyMethodDef gasp_methods[] = {
    {"run_committor_calc", (PyCFunction)(void (*)(void))run_committor_calc, METH_VARARGS | METH_KEYWORDS, comm_doc},
    {"run_nucleation_swarm", (PyCFunction)(void (*)(void))run_nucleation_swarm, METH_VARARGS | METH_KEYWORDS,
     nuc_doc},
    {nullptr, nullptr, 0, nullptr}};

static struct PyModuleDef gasp_module = {
    PyModuleDef_HEAD_INIT, "gasp",
    "Replica-parallel Metropolis engine for the 2D Ising model on AMD MI355X (PyGASP-compatible interface)", -1,
    gasp_methods, nullptr, nullptr, nullptr, nullptr};

PyMODINIT_FUNC PyInit_gasp(void) {
    import_array();
    GridSnapType.tp_name = "gasp.GridSnapObject";
    GridSnapType.tp_basicsize = sizeof(GridSnapObject);
    GridSnapType.tp_dealloc = (destructor)GridSnap_dealloc;
    GridSnapType.tp_flags = Py_TPFLAGS_DEFAULT | Py_TPFLAGS_BASETYPE;
    GridSnapType.tp_doc = "Grid snapshot: isweep, igrid, magnetisation, lclus_size, committor, grid (int8 LxL)";
    GridSnapType.tp_methods = GridSnap_methods;
    GridSnapType.tp_members = GridSnap_members;
    GridSnapType.tp_getset = GridSnap_getset;
    GridSnapType.tp_init = (initproc)GridSnap_init;
    GridSnapType.tp_new = PyType_GenericNew;
    if (PyType_Ready(&GridSnapType) < 0) return nullptr;

    PyObject *m = PyModule_Create(&gasp_module);
    if (!m) return nullptr;
    Py_INCREF(&GridSnapType);
    if (PyModule_AddObject(m, "GridSnapObject", (PyObject *)&GridSnapType) < 0) {
        Py_DECREF(&GridSnapType);
        Py_DECREF(m);
        return nullptr;
    }
    // src/ising_ext.cu:1179-1189: initialise the device at import
    g_idev = gpuInitDevice(-1, &g_gpu_nsms);
    if (g_idev < 0) {
        printf("Falling back to CPU\n");
        fflush(stdout);
        g_gpu_nsms = 0;
        gasp_clear_error();
    }
    g_module = m;
    if (PyModule_AddObject(m, "grids", PyList_New(0)) < 0 || PyModule_AddObject(m, "magnetisation", PyList_New(0)) < 0 ||
        PyModule_AddObject(m, "largest_cluster", PyList_New(0)) < 0 ||
        PyModule_AddObject(m, "gpu_nsms", PyLong_FromLong(g_gpu_nsms)) < 0) {
        Py_DECREF(m);
        return nullptr;
    }
    return m;
}
