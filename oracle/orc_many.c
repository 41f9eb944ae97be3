/*
 * orc_many.c -- bound many independent nodes with a pool of CPU threads.
 * Used only as the CPU baseline of the benchmark (one workspace per thread,
 * the reference's one-node-per-worker model without MPI).
 * TEST INFRASTRUCTURE ONLY.
 */
#include <stdlib.h>
#include <omp.h>

#include "orc_internal.h"

void orc_sdpbound_many(const double *SP_L, int n0, const orc_params *p, int count,
                       const int *xfixed, const int *solX, const double *LB,
                       const double *diff, const uint64_t *seeds, int nthreads, double *bounds,
                       int *status)
{
    int B = n0 - 1;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        orc_ws *w = orc_ws_create(n0, 400, 50000, 50000, 50000, p);
        double *frac = malloc(sizeof(double) * B);
        int *xb = malloc(sizeof(int) * B);
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < count; ++i) {
            orc_node_in in = {xfixed + (size_t)i * B, solX + (size_t)i * B, LB[i], diff ? diff[i] : 0.0, 0, seeds[i]};
            orc_node_out o;
            orc_sdpbound(w, SP_L, n0, &in, &o, frac, xb);
            bounds[i] = o.bound;
            status[i] = o.status;
        }
        free(frac);
        free(xb);
        orc_ws_destroy(w);
    }
}
