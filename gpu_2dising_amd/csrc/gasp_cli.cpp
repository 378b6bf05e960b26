// gasp_cli.cpp -- the `gasp` command line program (src/ising.cu:34-206).
//
//   gasp nsweeps nreplicas mag_output_int grid_output_int threadsPerBlock
//        gpu_device gpu_method beta h itask
//
// Same arguments, defaults, thresholds (+-0.90), output file (gridstates.bin,
// deleted at start) and stdout text as the reference.  Additions, all through
// the environment so that the argument list stays the reference's:
//   GASP_SEED   random seed (the reference seeds with time(NULL))
//   GASP_RNG    "philox" (default) or "mt19937" (per-replica MT19937 streams,
//               replica r seeded GASP_SEED + r: the reference CPU stream)
//   GASP_L      lattice size (the reference hard-codes 64)
// Without a visible GPU it falls back to the reference's CPU engine
// (mc_driver_cpu, one MT19937 stream seeded with the seed), as
// src/ising.cu:127-155 does; GASP_RNG then has no effect.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <vector>

#include "../../include/gasp_amd.h"

static int fail(const char *what) {
    fprintf(stderr, "gasp: %s failed (error %d: %s)\n", what, gasp_last_error(), gasp_error_string());
    return EXIT_FAILURE;
}

int main(int argc, char *argv[]) {
    int L = 64;
    const double dn_threshold = -0.90;
    const double up_threshold = 0.90;
    unsigned long rngseed = (unsigned long)time(NULL);

    if (argc != 11) {
        printf("Usage : GPU_2DIsing nsweeps nreplicas mag_output_int grid_output_int threadsPerBlock gpu_device "
               "gpu_method beta h itask \n");
        return EXIT_FAILURE;
    }
    const int tot_nsweeps = atoi(argv[1]);
    const int ngrids = atoi(argv[2]);
    const int mag_output_int = atoi(argv[3]);
    const int grid_output_int = atoi(argv[4]);
    const int threadsPerBlock = atoi(argv[5]);
    const int gpu_device = atoi(argv[6]);
    const int gpu_method = atoi(argv[7]);
    const double beta = atof(argv[8]);
    const double h = atof(argv[9]);
    const int itask = (int)atof(argv[10]);

    if (const char *s = getenv("GASP_SEED")) rngseed = strtoul(s, nullptr, 10);
    if (const char *s = getenv("GASP_L")) L = atoi(s);
    bool use_mt = false;
    if (const char *s = getenv("GASP_RNG")) use_mt = strcmp(s, "mt19937") == 0;

    remove("gridstates.bin");
    if (itask == 0) printf("# isweep    nucleated fraction\n");

    int *ising_grids = nullptr;
    int *grid_fate = nullptr;
    if (itask == 0) {
        ising_grids = init_grids_uniform(L, ngrids, -1);
        if (!ising_grids) return fail("init_grids_uniform");
    } else if (itask == 1) {
        ising_grids = init_grids_from_file(L, ngrids);
        if (!ising_grids) return fail("init_grids_from_file");
        grid_fate = init_fates(ngrids);
        if (!grid_fate) return fail("init_fates");
    } else {
        printf("Error - unknown value of itask!");
        return EXIT_FAILURE;
    }

    // the reference passes a single float here although itask 0 fills
    // tot_nsweeps/mag_output_int values
    const int nres = itask == 0 ? (mag_output_int > 0 ? tot_nsweeps / mag_output_int : 0) : 1;
    std::vector<float> result(nres > 0 ? nres : 1, 0.0f);
    char cv[] = "magnetisation";
    char fname[] = "gridstates.bin";
    mc_sampler_t samples;
    samples.tot_nsweeps = tot_nsweeps;
    samples.mag_output_int = mag_output_int;
    samples.grid_output_int = grid_output_int;
    mc_function_t calc;
    calc.itask = itask;
    calc.dn_thr = dn_threshold;
    calc.up_thr = up_threshold;
    calc.ninputs = 1;
    calc.initial_spin = -1;
    calc.cv = cv;
    calc.result = result.data();
    calc.filename = fname;

    int gpu_nsms = 0;
    if (gpuInitDevice(gpu_device, &gpu_nsms) == -1) {
        // src/ising.cu:131-155
        printf("Falling back to CPU\n");
        fflush(stdout);
        gasp_clear_error();
        init_genrand(rngseed);
        preComputeProbs_cpu(beta, h);
        mc_grids_t cgrids;
        cgrids.L = L;
        cgrids.ngrids = ngrids;
        cgrids.ising_grids = ising_grids;
        mc_driver_cpu(cgrids, beta, h, grid_fate, samples, calc, write_ising_grids);
        const int rc = gasp_last_error();
        free(ising_grids);
        free(grid_fate);
        return rc ? EXIT_FAILURE : EXIT_SUCCESS;
    }

    int *d_ising_grids = nullptr, *d_neighbour_list = nullptr;
    gpuInitGrid(L, ngrids, threadsPerBlock, ising_grids, &d_ising_grids, &d_neighbour_list);
    if (gasp_last_error()) return fail("gpuInitGrid");

    gasp_rng_state *d_state = nullptr;
    if (use_mt) {
        if (gasp_init_rand_mt(ngrids, rngseed, &d_state)) return fail("gasp_init_rand_mt");
    } else {
        gpuInitRand(ngrids, threadsPerBlock, rngseed, &d_state);
        if (gasp_last_error()) return fail("gpuInitRand");
    }
    preComputeProbs_gpu(beta, h);

    mc_gpu_grids_t grids;
    grids.L = L;
    grids.ngrids = ngrids;
    grids.ising_grids = ising_grids;
    grids.d_ising_grids = d_ising_grids;
    grids.d_neighbour_list = d_neighbour_list;
    gpu_run_t gpu_state;
    gpu_state.d_state = d_state;
    gpu_state.threadsPerBlock = threadsPerBlock;
    gpu_state.gpu_method = gpu_method;

    mc_driver_gpu(grids, beta, h, grid_fate, samples, calc, gpu_state, write_ising_grids);
    const int rc = gasp_last_error();

    gasp_free(d_state);
    gasp_free(d_ising_grids);
    gasp_free(d_neighbour_list);
    free(ising_grids);
    free(grid_fate);
    return rc ? EXIT_FAILURE : EXIT_SUCCESS;
}
