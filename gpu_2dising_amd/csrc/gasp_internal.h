// gasp_internal.h -- declarations shared by the engine's host translation units.
#pragma once
#include <stdarg.h>
#include <stdint.h>

#include "../../include/gasp_amd.h"

namespace gasp {

// Records an error (code + message), prints the message to stderr like the
// reference's fprintf(stderr, ...) + exit paths, and returns the code.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Text output of the driver.  Lines go to stdout (and are flushed) unless a
// callback is installed (the Python module routes them to sys.stdout).
void emit(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Output style of the driver's progress lines: 0 = command line build
// (src/mc_gpu.cu '#ifndef PYTHON' branches), 1 = Python build ('#ifdef PYTHON').
int output_style();

// Exact integer threshold T of "(u + 0.5) 2^-32 < p" (u < T), T in [0, 2^32].
uint64_t threshold_of(double p);

// Appends one gridstates.bin record (int32 L, ngrids, isweep + the packed
// bits, src/io.c:67-130) to `filename` (NULL: gridstates.bin).
int append_gridstates(int L, int ngrids, int isweep, const uint8_t *bits, const char *filename);

}  // namespace gasp

extern "C" {
// Routes the driver's text output (NULL restores stdout) and selects its style.
void gasp_set_output(void (*cb)(const char *text), int style);
}
