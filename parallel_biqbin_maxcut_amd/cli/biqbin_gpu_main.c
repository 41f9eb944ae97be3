/*
 * biqbin_gpu -- command-line solver: `biqbin_gpu [options] graph.rudy params
 * [batch]`, the reference's `./biqbin file.rudy file.params` (src/main.c,
 * src/process_input.c:37-99) on one GPU.  Prints the reference's final
 * report (src/bab_functions.c:324-367): nodes, root bound, maximum value,
 * solution (1-based vertices on side 1) and time.
 *
 * Options (bqb_solve_options):
 *   -w          children inherit their parent's final cuts and multipliers
 *   -p auto     the size class's precision from the measured table, with
 *               certified bounds
 *   -p N        IPM products on the bf16 matrix cores with N slices (2..4),
 *               certified bounds
 *   -c          certified bounds with FP64 products
 *   -g          launch trains as HIP graphs
 *   -i          bundle-step IPMs start from the node's previous solve
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "biqbin_gpu.h"

static int usage(const char *prog)
{
    fprintf(stderr, "Usage: %s [-w] [-p auto|2|3|4] [-c] [-g] [-i] file.rudy file.params [nodes per batch]\n", prog);
    return 1;
}

int main(int argc, char **argv)
{
    bqb_solve_options o = {.batch = 64, .max_bundle = 0, .slices = 0, .certify = 0, .model = NULL,
                           .drop = 0.f, .warm = 0, .graphs = -1, .ipm_warm = 0};
    int opt;
    while ((opt = getopt(argc, argv, "wp:cgi")) != -1) {
        switch (opt) {
        case 'w': o.warm = 1; break;
        case 'p':
            if (strcmp(optarg, "auto") == 0) o.slices = -1;
            else {
                o.slices = atoi(optarg);
                if (o.slices < 2 || o.slices > 4) return usage(argv[0]);
            }
            break;
        case 'c': o.certify = 1; break;
        case 'g': o.graphs = 1; break;
        case 'i': o.ipm_warm = 1; break;
        default: return usage(argv[0]);
        }
    }
    const int npos = argc - optind;
    if (npos < 2 || npos > 3) return usage(argv[0]);
    const char *graph = argv[optind], *params = argv[optind + 1];
    if (npos == 3) o.batch = atoi(argv[optind + 2]);
    BiqBinParameters p;
    if (bqb_read_params(params, &p)) { fprintf(stderr, "Error: %s\n", bqb_last_error()); return 1; }
    int nv = 0;
    double *adj = bqb_read_rudy(graph, &nv);
    if (!adj) { fprintf(stderr, "Error: %s\n", bqb_last_error()); return 1; }
    printf("Input file: %s\n\nGraph has %d vertices.\n", graph, nv);
    double *L = malloc(sizeof(double) * (size_t)nv * nv);
    int *sol = calloc(nv, sizeof(int));
    if (!L || !sol || bqb_build_L(nv, adj, L)) { fprintf(stderr, "Error: %s\n", bqb_last_error()); return 1; }
    free(adj);
    if (p.adjust_TriIneq) p.TriIneq = nv * 10;
    bqb_solve_stats st;
    const int rc = bqb_solve_opt(nv, L, &p, 2020, &o, sol, &st);
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
           st.bound_ms / 1000.0, st.batches, o.batch);
    free(L);
    free(sol);
    return 0;
}
