/* bqb_phase.h -- phases of the batched SDPbound control kernel. */
#ifndef BQB_PHASE_H
#define BQB_PHASE_H
enum {
    PH_SETUP = 0,        /* createSubproblem, getFixedValue, node state (stage NEW) */
    PH_POST_BASIC = 1,   /* f, basic bound, fracsol after the basic IPM */
    PH_CHECK_BASIC = 2,  /* prune / give-up after the first heuristic */
    PH_BUNDLE_INIT = 3,  /* bundle initialisation after the first separation */
    PH_OUTER_BEGIN = 4,  /* ++count, oldf */
    PH_OUTER_A = 5,      /* bound and prune after the bundle method */
    PH_OUTER_B = 6,      /* prune / give-up after the heuristic */
    PH_SEP_PENT = 7,     /* pentagonal separation decision */
    PH_SEP_HEPT = 8,     /* heptagonal separation decision */
    PH_OUTER_C = 9,      /* stop test, fracsol, gamma from the cuts */
    PH_REBUILD = 10,     /* extend the bundle after fct_eval(gamma) */
    PH_FINISH = 11,
    PH_FCT_FINISH = 12,  /* fct_eval tail: f = phi + sum(gamma), g = 1 - B(X_test) */
    PH_LOAD = 13,        /* copy staged nodes into their slots, mask = newly loaded */
    PH_BUNDLE_NEXT = 14, /* mask for bundle iteration arg: active LOOP nodes with arg < bdl_iter */
    PH_WARM = 15,        /* mask: nodes whose inherited multipliers are evaluated next */
    PH_CERT_MASK = 16    /* mask: nodes at a round end whose center moved since its certificate */
};
#endif
