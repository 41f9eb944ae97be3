/*
 * biqbin_gpu -- command-line solver: `biqbin_gpu graph.rudy params [batch]`,
 * the reference's `./biqbin file.rudy file.params` (src/main.c,
 * src/process_input.c:37-99) on one GPU.  Prints the reference's final
 * report (src/bab_functions.c:324-367): nodes, root bound, maximum value,
 * solution (1-based vertices on side 1) and time.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "biqbin_gpu.h"

int main(int argc, char **argv)
{
    if (argc < 3 || argc > 4) {
        fprintf(stderr, "Usage: %s file.rudy file.params [nodes per batch]\n", argv[0]);
        return 1;
    }
    const int batch = argc == 4 ? atoi(argv[3]) : 64;
    BiqBinParameters p;
    if (bqb_read_params(argv[2], &p)) { fprintf(stderr, "Error: %s\n", bqb_last_error()); return 1; }
    int nv = 0;
    double *adj = bqb_read_rudy(argv[1], &nv);
    if (!adj) { fprintf(stderr, "Error: %s\n", bqb_last_error()); return 1; }
    printf("Input file: %s\n\nGraph has %d vertices.\n", argv[1], nv);
    double *L = malloc(sizeof(double) * (size_t)nv * nv);
    int *sol = calloc(nv, sizeof(int));
    if (!L || !sol || bqb_build_L(nv, adj, L)) { fprintf(stderr, "Error: %s\n", bqb_last_error()); return 1; }
    free(adj);
    if (p.adjust_TriIneq) p.TriIneq = nv * 10;
    bqb_solve_stats st;
    const int rc = bqb_solve(nv, L, &p, 2020, batch, sol, &st);
    if (rc) { fprintf(stderr, "Error: %s\n", bqb_last_error()); return 10; }
    printf("\nNodes = %lld\n", st.nodes);
    printf("Root node bound = %.2lf\n", st.root_bound);
    printf("%s = %.0lf\n", st.stopped ? "Best value" : "Maximum value", st.best);
    if (st.stopped) printf("TIME LIMIT REACHED.\n");
    printf("Solution = ( ");
    for (int i = 0; i < nv - 1; ++i)
        if (sol[i] == 1) printf("%d ", i + 1);
    printf(")\n");
    printf("Time = %.2f s (bounding %.2f s, %lld batches of up to %d nodes)\n\n", st.seconds,
           st.bound_ms / 1000.0, st.batches, batch);
    free(L);
    free(sol);
    return 0;
}
