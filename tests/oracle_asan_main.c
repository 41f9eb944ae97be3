/* Driver for the sanitizer run of the CPU oracle: a full branch and bound
 * with all cut types on a small random graph, and a node bound with
 * fixings, built with AddressSanitizer and UndefinedBehaviorSanitizer. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int main(void)
{
    const int n0 = 26;
    double *adj = calloc(n0 * n0, sizeof(double)), *L = calloc(n0 * n0, sizeof(double));
    uint64_t s = 12345;
    for (int i = 0; i < n0; ++i)
        for (int j = i + 1; j < n0; ++j) {
            s = orc_mix64(s + 0x9E3779B97F4A7C15ull);
            if ((s >> 11) % 2) {
                const double w = (double)((int)((s >> 20) % 21) - 10);
                adj[i + j * n0] = adj[j + i * n0] = w;
            }
        }
    orc_build_L_from_adj(adj, n0, L);
    orc_params p;
    orc_params_default(&p);
    p.TriIneq = 10 * n0;
    int *sol = calloc(n0, sizeof(int));
    orc_bab_out out;
    orc_bab_solve(L, n0, &p, 7, 4, sol, &out);
    printf("best %.1f nodes %d status %d\n", out.best, out.nodes, out.status);
    int *xf = calloc(n0, sizeof(int)), *sx = calloc(n0, sizeof(int));
    xf[3] = xf[7] = xf[11] = 1;
    sx[7] = 1;
    orc_node_in in = {xf, sx, 0.0, 0.0, 0, 99};
    orc_node_out no;
    double *frac = calloc(n0, sizeof(double));
    int *xb = calloc(n0, sizeof(int));
    orc_ws *w = orc_ws_create(n0, 400, 5000, 5000, 5000, &p);
    orc_sdpbound(w, L, n0, &in, &no, frac, xb);
    orc_ws_destroy(w);
    printf("node bound %.6f status %d\n", no.bound, no.status);
    free(adj); free(L); free(sol); free(xf); free(sx); free(frac); free(xb);
    return (out.status || no.status) ? 1 : 0;
}
