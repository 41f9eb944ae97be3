// bqb_control.hip -- per-node control phases of the batched SDPbound
// (reference src/bounding.c:24-342 and src/evaluate.c:11-138), plus the
// cut-operator kernels that the bundle loop needs:
//   - subproblem construction (createSubproblem / getFixedValue),
//   - the prune / give-up / separation decisions of the outer loop,
//   - the op_Bt gather lists and L0 = L - B'(gamma) (operators.c:140-481),
//   - bundle initialisation and the rebuild after a cut change.
// One workgroup per node slot; nodes whose phase mask is off return at once.
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_cutops.h"
#include "bqb_kernels.h"
#include "bqb_phase.h"

namespace bqb {

constexpr int CT = 256;   // threads of a control workgroup

__device__ __forceinline__ double block_sum(double v, double *red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < CT / 64; ++w) s += red[w];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double block_max(double v, double *red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < CT / 64; ++w) s = fmax(s, red[w]);
    __syncthreads();
    return s;
}

// <L, X> over the leading n x n block.
__device__ double dot_LX(const double *L, const double *X, int n, int ld, double *red)
{
    double s = 0.0;
    for (int e = threadIdx.x; e < n * n; e += CT) {
        const int i = e % n, j = e / n;
        s += L[i + (size_t)j * ld] * X[i + (size_t)j * ld];
    }
    return block_sum(s, red);
}

// fracsol from the last column of X (bounding.c:79-89, 252-262)
__device__ void store_fracsol(const BatchDims &d, const BatchPtrs &p, int node, const NodeState &s)
{
    const int B = d.n0 - 1;
    const int *xf = p.xfixed + (size_t)node * B;
    const int *sx = p.solX + (size_t)node * B;
    const double *X = p.X + (size_t)node * d.ld * d.ld;
    double *fr = p.fracsol + (size_t)node * B;
    if (threadIdx.x == 0) {
        int index = 0;
        for (int i = 0; i < B; ++i) {
            if (xf[i]) fr[i] = (double)sx[i];
            else { fr[i] = 0.5 * (X[(s.n - 1) + (size_t)index * d.ld] + 1.0); ++index; }
        }
    }
}

// g = 1 - B(X) for the node's m cuts
__device__ void subgradient(const BatchDims &d, const BatchPtrs &p, int node, const NodeState &s,
                            const double *X, double *g)
{
    const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
    const TriCut *tri = p.tri + (size_t)node * d.max_tri;
    const KgonCut *pent = p.pent + (size_t)node * d.max_pent;
    const KgonCut *hept = p.hept + (size_t)node * d.max_hept;
    for (int c = threadIdx.x; c < m; c += CT)
        g[c] = 1.0 - cut_B(X, d.ld, c, s.NIneq, s.NPentIneq, tri, pent, hept);
}

__global__ __launch_bounds__(CT) void control_kernel(BatchDims d, BatchPtrs p, int phase, int arg)
{
    __shared__ double red[CT / 64];
    __shared__ int s_free[1024];
    __shared__ double s_fr[1024];
    __shared__ int s_nf;
    const int node = blockIdx.x;
    if (node >= d.nodes) return;
    NodeState *sp = p.st + node;
    const BqbParams P = *p.params;
    const int B = d.n0 - 1;
    const size_t slab = (size_t)d.ld * d.ld;

    if (phase == PH_LOAD) {
        // staged node -> slot; every slot's mask says whether it was loaded
        const int k = p.ld_of_slot[node];
        if (k >= 0) {
            if (threadIdx.x == 0) *sp = p.ld_st[k];
            for (int i = threadIdx.x; i < B; i += CT) {
                const_cast<int *>(p.xfixed)[(size_t)node * B + i] = p.ld_xf[(size_t)k * B + i];
                const_cast<int *>(p.solX)[(size_t)node * B + i] = p.ld_sx[(size_t)k * B + i];
            }
        }
        if (threadIdx.x == 0) p.mask[node] = (k >= 0) ? 1 : 0;
        return;
    }

    if (phase == PH_SETUP) {
        // createSubproblem (evaluate.c:38-115) into the zero-padded slab,
        // getFixedValue (evaluate.c:123-138).  Integer-valued data: every
        // sum below is exact, so the parallel order does not matter.
        if (sp->stage != BQB_STAGE_NEW || !p.mask[node]) return;
        const int *xf = p.xfixed + (size_t)node * B;
        const int *sx = p.solX + (size_t)node * B;
        double *L = p.L + node * slab;
        if (threadIdx.x == 0) {
            int nf = 0;
            for (int i = 0; i < B; ++i)
                if (!xf[i]) s_free[nf++] = i;
            s_nf = nf;
        }
        __syncthreads();
        const int nf = s_nf;
        const int n = nf + 1;
        const int N = d.n0;
        for (int e = threadIdx.x; e < (int)slab; e += CT) L[e] = 0.0;
        for (int a = threadIdx.x; a < nf; a += CT) {
            const int j = s_free[a];
            double fr = 0.0;
            for (int i = 0; i < B; ++i)
                if (xf[i] && sx[i] == 1) fr += p.SP_L[j + (size_t)i * N];
            s_fr[a] = fr;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nf * nf; e += CT) {
            const int b = e % nf, a = e / nf;
            double v = p.SP_L[s_free[b] + (size_t)s_free[a] * N];
            if (a == b) v += 2 * s_fr[a];
            L[b + (size_t)a * d.ld] = v;
        }
        for (int a = threadIdx.x; a < nf; a += CT) {
            double rs = 0.0;
            for (int b = 0; b < nf; ++b) rs += p.SP_L[s_free[b] + (size_t)s_free[a] * N];
            rs += 2 * s_fr[a];
            L[(n - 1) + (size_t)a * d.ld] = rs;
            L[a + (size_t)(n - 1) * d.ld] = rs;
        }
        __syncthreads();
        double cs = 0.0;
        for (int a = threadIdx.x; a < nf; a += CT) cs += L[a + (size_t)(n - 1) * d.ld];
        const double corner = block_sum(cs, red);
        double fv = 0.0;
        for (int e = threadIdx.x; e < B * B; e += CT) {
            const int i = e / B, j = e % B;
            if (xf[i] && xf[j]) fv += p.SP_L[j + (size_t)i * N] * sx[i] * sx[j];
        }
        fv = block_sum(fv, red);
        if (threadIdx.x == 0) L[(n - 1) + (size_t)(n - 1) * d.ld] = corner;
        __syncthreads();
        for (int e = threadIdx.x; e < n * n; e += CT) {
            const int i = e % n, j = e / n;
            L[i + (size_t)j * d.ld] *= 0.25;
        }
        // A node loaded with inherited cuts (records and multipliers already
        // in its cut arrays) keeps them: the multipliers become the current
        // gamma, laid out [triangles | pentagonal | heptagonal], which the
        // first triangle update reads for its purge.
        const int warm = sp->warm;
        const int wt = warm ? sp->NIneq : 0, wp = warm ? sp->NPentIneq : 0, wh = warm ? sp->NHeptaIneq : 0;
        if (warm) {
            double *gam = p.gamma + (size_t)node * d.max_m;
            for (int c = threadIdx.x; c < wt; c += CT) gam[c] = p.tri_y[(size_t)node * d.max_tri + c];
            for (int c = threadIdx.x; c < wp; c += CT) gam[wt + c] = p.pent_y[(size_t)node * d.max_pent + c];
            for (int c = threadIdx.x; c < wh; c += CT) gam[wt + wp + c] = p.hept_y[(size_t)node * d.max_hept + c];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            sp->n = n;
            sp->fixedvalue = fv;
            sp->status = 0;
            sp->active = 1;
            sp->prune = sp->giveup = sp->done = 0;
            sp->count = 0;
            sp->bdl_iter = P.init_bundle_iter;
            sp->NIneq = wt;
            sp->NPentIneq = wp;
            sp->NHeptaIneq = wh;
            sp->bundle = 0;
            sp->heur_idx = 0;
            sp->improved = 0;
            sp->ipm_calls = sp->ipm_iters = 0;
            sp->max_cuts = sp->max_bundle_seen = 0;
            sp->viol3 = sp->viol5 = 0.0;
            sp->f_cert = INFINITY;       // until a certificate is computed
            sp->cert_stale = 0;
            sp->cuts_changed = 1;
            p.nvec[node] = n;
            p.mask[node] = 1;
        }
        return;
    }

    NodeState s = *sp;   // snapshot (written back by thread 0 where changed)
    __syncthreads();

    // Each phase acts on the nodes of one stage; every other slot has its
    // mask cleared so the kernels that follow skip it.
    {
        int want = -1;
        switch (phase) {
        case PH_POST_BASIC: case PH_CHECK_BASIC: case PH_BUNDLE_INIT: want = BQB_STAGE_NEW; break;
        case PH_OUTER_BEGIN: case PH_OUTER_A: case PH_OUTER_B: case PH_SEP_PENT: case PH_SEP_HEPT:
        case PH_OUTER_C: case PH_REBUILD: case PH_BUNDLE_NEXT: case PH_WARM: case PH_CERT_MASK:
            want = BQB_STAGE_LOOP; break;
        default: break;
        }
        // outer rounds advance per node: begin (rphase 0), bundle
        // iterations (1), round end (2)
        int want_r = -1;
        switch (phase) {
        case PH_OUTER_BEGIN: want_r = 0; break;
        case PH_BUNDLE_NEXT: want_r = 1; break;
        case PH_OUTER_A: case PH_OUTER_B: case PH_SEP_PENT: case PH_SEP_HEPT: case PH_OUTER_C:
        case PH_REBUILD: case PH_WARM: case PH_CERT_MASK: want_r = 2; break;
        default: break;
        }
        if ((want >= 0 && s.stage != want) || (want_r >= 0 && s.rphase != want_r)) {
            if (threadIdx.x == 0) p.mask[node] = 0;
            return;
        }
    }

    if (phase == PH_BUNDLE_NEXT) {
        if (threadIdx.x == 0) p.mask[node] = (s.active && s.bdl_left > 0) ? 1 : 0;
        return;
    }

    if (phase == PH_WARM) {
        if (threadIdx.x == 0) p.mask[node] = (s.active && s.warm) ? 1 : 0;
        return;
    }

    if (phase == PH_CERT_MASK) {
        if (threadIdx.x == 0) p.mask[node] = (s.active && s.cert_stale) ? 1 : 0;
        return;
    }

    if (phase == PH_POST_BASIC) {
        // bounding.c:73-103: f from the basic IPM, fracsol, heuristic mask
        if (!s.active) { if (threadIdx.x == 0) p.mask[node] = 0; return; }
        if (threadIdx.x == 0) {
            sp->f = p.phi[node];
            sp->basic_bound = sp->f + s.fixedvalue;
            sp->heur_idx = 0;
            sp->ipm_calls += 1;
            sp->ipm_iters += p.ipm_iters[node];
            if (p.ipm_status[node]) { sp->status = p.ipm_status[node]; sp->active = 0; }
            sp->bound = sp->basic_bound;
            p.mask[node] = sp->active;
        }
        s.n = sp->n;
        store_fracsol(d, p, node, s);
        return;
    }

    if (phase == PH_CHECK_BASIC) {
        // bounding.c:106-121: prune / give up, then first triangle separation
        if (threadIdx.x == 0) {
            int m = 0;
            if (s.active && !s.status) {
                // the prune test reads the certified bound where bounds are
                // certified; the give-up test (like every choice that only
                // steers the work) reads f + fixedvalue, as the reference
                const double bound = node_bound(s, d.certify);
                sp->bound = bound;
                if (bound < s.LB + 1.0) { sp->prune = 1; sp->active = 0; }
                else if (P.use_diff && !s.is_root && s.f + s.fixedvalue > s.LB + s.diff + 1.0) {
                    sp->giveup = 1;
                    sp->active = 0;
                }
                else m = 1;
            } else sp->active = 0;
            sp->sep_tri = m;
            p.mask[node] = m;
        }
        return;
    }

    if (phase == PH_BUNDLE_INIT) {
        // bounding.c:127-163
        if (!s.active) { if (threadIdx.x == 0) p.mask[node] = 0; return; }
        // m = NIneq unless the node inherited pentagonal / heptagonal cuts
        const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
        double *gam = p.gamma + (size_t)node * d.max_m;
        double *g = p.g + (size_t)node * d.max_m;
        double *G = p.G + (size_t)node * d.max_bundle * d.max_m;
        const double *X = p.X + node * slab;
        for (int c = threadIdx.x; c < s.NIneq; c += CT) gam[c] = p.tri_y[(size_t)node * d.max_tri + c];
        for (int c = threadIdx.x; c < s.NPentIneq; c += CT) gam[s.NIneq + c] = p.pent_y[(size_t)node * d.max_pent + c];
        for (int c = threadIdx.x; c < s.NHeptaIneq; c += CT)
            gam[s.NIneq + s.NPentIneq + c] = p.hept_y[(size_t)node * d.max_hept + c];
        subgradient(d, p, node, s, X, g);
        __syncthreads();
        for (int c = threadIdx.x; c < m; c += CT) G[c] = g[c];
        double *Xb = p.Xb + (size_t)node * d.max_bundle * slab;
        for (InteriorWalk w(s.n, threadIdx.x, CT); w.more(); w.next()) {
            const size_t e = w.off(d.ld);
            Xb[e] = X[e];
        }
        const double F0 = dot_LX(p.L + node * slab, X, s.n, d.ld, red);
        if (threadIdx.x == 0) {
            p.F[(size_t)node * d.max_bundle] = F0;
            sp->bundle = 1;
            sp->max_bundle_seen = 1;
            sp->t = 0.5 * (s.f + s.fixedvalue - s.LB) / (m * s.viol3 * s.viol3);
            sp->cuts_changed = 1;
            sp->stage = BQB_STAGE_LOOP;
            // an inherited node evaluates its multipliers (PH_WARM, op_Bt,
            // IPM, PH_REBUILD) before its first outer round
            sp->rphase = s.warm ? 2 : 0;
            p.mask[node] = 1;
        }
        return;
    }

    if (phase == PH_OUTER_BEGIN) {
        // bounding.c:170-174; the bundle length of outer round c >= 2 is
        // init_bundle_iter + floor(c/2), at most max_bundle_iter
        // (bounding.c:323-325 add c % 2 after every round)
        if (threadIdx.x == 0) {
            if (s.active) {
                const int c = s.count + 1;
                const int bdl = P.init_bundle_iter + c / 2;
                sp->count = c;
                sp->oldf = s.f;
                // round 1 runs init_bundle_iter as given; the cap applies
                // from the first increment on (bounding.c:50, 323-325)
                sp->bdl_iter = (c == 1) ? P.init_bundle_iter
                                        : (bdl < P.max_bundle_iter ? bdl : P.max_bundle_iter);
                sp->bdl_left = sp->bdl_iter;
                sp->rphase = sp->bdl_left > 0 ? 1 : 2;
            }
            p.mask[node] = s.active;
        }
        return;
    }

    if (phase == PH_OUTER_A) {
        // bounding.c:177-183: bound after the bundle, heuristic unless pruned
        if (threadIdx.x == 0) {
            int h = 0;
            if (s.active) {
                const double bound = node_bound(s, d.certify);
                sp->bound = bound;
                sp->prune = (bound < s.LB + 1.0) ? 1 : 0;
                sp->heur_idx = s.count;
                h = !sp->prune;
            }
            p.mask[node] = h;
        }
        return;
    }

    if (phase == PH_OUTER_B) {
        // bounding.c:197-222: prune after the heuristic, give-up tests,
        // separation decision
        if (threadIdx.x == 0) {
            int sep = 0;
            if (s.active) {
                int prune = s.prune;
                if (!prune) prune = (s.bound < s.LB + 1.0) ? 1 : 0;
                const double gap = s.f + s.fixedvalue - s.LB;   // steers only: uncertified
                int giveup = 0;
                if (s.count == P.triag_iter + P.pent_iter + P.hept_iter)
                    if (gap - 1.0 > (s.oldf - s.f) * (P.max_outer_iter - s.count)) giveup = 1;
                if (s.count == P.max_outer_iter)
                    if (gap - 1.0 > (s.oldf - s.f) * P.extra_iter) giveup = 1;
                if (s.count == P.max_outer_iter + P.extra_iter) giveup = 1;
                sp->prune = prune;
                sp->giveup = giveup;
                sp->gap = gap;
                sep = !prune && !giveup;
            }
            sp->sep_tri = sep;
            sp->sep_pent = 0;
            sp->sep_hept = 0;
            p.mask[node] = sep;
        }
        return;
    }

    if (phase == PH_SEP_PENT || phase == PH_SEP_HEPT) {
        // bounding.c:231-235
        if (threadIdx.x == 0) {
            int on = 0;
            if (s.active && s.sep_tri) {
                if (phase == PH_SEP_PENT)
                    on = P.include_Pent && (s.count > P.triag_iter || s.viol3 < 0.2);
                else
                    on = P.include_Hepta && ((s.count > P.triag_iter + P.pent_iter) ||
                                             (s.viol3 < 0.2 && (1 - s.viol5 < 0.4)));
            }
            if (phase == PH_SEP_PENT) sp->sep_pent = on; else sp->sep_hept = on;
            p.mask[node] = on;
        }
        return;
    }

    if (phase == PH_OUTER_C) {
        // bounding.c:247-275: stop test, fracsol, gamma from the cut lists
        if (!s.active) { if (threadIdx.x == 0) p.mask[node] = 0; return; }
        const int done = s.prune || s.giveup || s.status;
        store_fracsol(d, p, node, s);
        if (done) {
            if (threadIdx.x == 0) {
                sp->done = 1;
                sp->active = 0;
                sp->bound = node_bound(s, d.certify);
                if (s.is_root && !s.status) sp->diff = s.basic_bound - (s.f + s.fixedvalue);
                p.mask[node] = 0;
            }
            return;
        }
        double *gam = p.gamma + (size_t)node * d.max_m;
        for (int c = threadIdx.x; c < s.NIneq; c += CT) gam[c] = p.tri_y[(size_t)node * d.max_tri + c];
        for (int c = threadIdx.x; c < s.NPentIneq; c += CT) gam[s.NIneq + c] = p.pent_y[(size_t)node * d.max_pent + c];
        for (int c = threadIdx.x; c < s.NHeptaIneq; c += CT)
            gam[s.NIneq + s.NPentIneq + c] = p.hept_y[(size_t)node * d.max_hept + c];
        if (threadIdx.x == 0) {
            const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
            if (m > s.max_cuts) sp->max_cuts = m;
            sp->cuts_changed = 1;
            p.mask[node] = 1;
        }
        return;
    }

    if (phase == PH_REBUILD) {
        // bounding.c:278-321: extend the bundle with fct_eval(gamma) after a
        // cut change; G is recomputed for the new cut set.
        // arg 1: the evaluation of inherited multipliers (warm nodes only)
        if (!s.active || (arg == 1 && !s.warm)) { if (threadIdx.x == 0) p.mask[node] = 0; return; }
        const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
        if (s.bundle >= d.max_bundle) {
            if (threadIdx.x == 0) bundle_full_giveup(sp, s, p.mask + node, d.certify);
            return;
        }
        double *G = p.G + (size_t)node * d.max_bundle * d.max_m;
        const double *Xb = p.Xb + (size_t)node * d.max_bundle * slab;
        const double *Xt = p.Xt + node * slab;
        const TriCut *tri = p.tri + (size_t)node * d.max_tri;
        const KgonCut *pent = p.pent + (size_t)node * d.max_pent;
        const KgonCut *hept = p.hept + (size_t)node * d.max_hept;
        for (int e = threadIdx.x; e < m * s.bundle; e += CT) {
            const int c = e % m, i = e / m;
            G[e] = 1.0 - cut_B(Xb + i * slab, d.ld, c, s.NIneq, s.NPentIneq, tri, pent, hept);
        }
        double *g = p.g + (size_t)node * d.max_m;
        subgradient(d, p, node, s, Xt, g);
        __syncthreads();
        for (int c = threadIdx.x; c < m; c += CT) G[(size_t)s.bundle * m + c] = g[c];
        double *Xdst = p.Xb + ((size_t)node * d.max_bundle + s.bundle) * slab;
        for (InteriorWalk w(s.n, threadIdx.x, CT); w.more(); w.next()) {
            const size_t e = w.off(d.ld);
            Xdst[e] = Xt[e];
        }
        const double Fb = dot_LX(p.L + node * slab, Xt, s.n, d.ld, red);
        // With warm-started multipliers the new stability center gamma is no
        // longer (nearly) the old one, so its value f(gamma) = phi + sum(gamma)
        // from this evaluation becomes f (any gamma >= 0 gives a valid bound).
        // An inherited node's first center: f(gamma) of the inherited
        // multipliers if it is below the basic value f(0), else gamma = 0 as
        // in the reference (the evaluation stays in the bundle either way).
        double sg = 0.0;
        double *gam = p.gamma + (size_t)node * d.max_m;
        if (d.learned || s.warm) {
            for (int c = threadIdx.x; c < m; c += CT) sg += gam[c];
            sg = block_sum(sg, red);
        }
        const double fnew = p.phi[node] + sg;
        const bool keep_center = !s.warm || (fnew < s.f && !p.ipm_status[node]);
        if (!keep_center)
            for (int c = threadIdx.x; c < m; c += CT) gam[c] = 0.0;
        // The learned warm start of the new cuts moves the center to them
        // only where it lowers f; otherwise the new cuts start at 0 as in the
        // reference (cutting_planes.c:201, 392, 588) and f stays (bounding.c:
        // 278 extends the bundle without a new center).
        const bool learned_keep = !d.learned || s.warm || (fnew < s.f && !p.ipm_status[node]);
        if (!learned_keep) {
            const int a0 = s.nsurv_tri, a1 = s.NIneq + s.nsurv_pent, a2 = s.NIneq + s.NPentIneq + s.nsurv_hept;
            for (int c = threadIdx.x; c < m; c += CT) {
                const bool fresh = (c >= a0 && c < s.NIneq) || (c >= a1 && c < s.NIneq + s.NPentIneq) || c >= a2;
                if (fresh) gam[c] = 0.0;
            }
        }
        // a center moved here (warm-started multipliers) is certified at the
        // end of the round like one moved by a serious step: its y is kept,
        // and until then the last certificate stays a valid bound
        const bool moved = d.certify && !p.ipm_status[node] && (s.warm ? keep_center : d.learned && learned_keep);
        if (moved) {
            const double *yt = p.ytest + (size_t)node * d.ld;
            double *yc = p.ycen + (size_t)node * d.ld;
            for (int i = threadIdx.x; i < s.n; i += CT) yc[i] = yt[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const int failed = p.ipm_status[node];
            sp->ipm_calls += 1;
            sp->ipm_iters += p.ipm_iters[node];
            sp->rphase = 0;
            if (moved) sp->cert_stale = 1;
            if (s.warm) {
                // a failed evaluation of inherited multipliers adds nothing
                // and leaves the node on its reference path (gamma = 0)
                if (failed) sp->flags = s.flags | BQB_FLAG_WARM_FAILED;
                else {
                    if (keep_center) sp->f = fnew;
                    p.F[(size_t)node * d.max_bundle + s.bundle] = Fb;
                    sp->bundle = s.bundle + 1;
                    if (sp->bundle > s.max_bundle_seen) sp->max_bundle_seen = sp->bundle;
                }
                sp->warm = 0;
                p.mask[node] = keep_center ? 1 : 0;
            } else {
                if (d.learned && learned_keep) sp->f = fnew;
                p.F[(size_t)node * d.max_bundle + s.bundle] = Fb;
                sp->bundle = s.bundle + 1;
                if (sp->bundle > s.max_bundle_seen) sp->max_bundle_seen = sp->bundle;
                sp->t = s.t * 1.05;
                if (failed) { sp->status = failed; sp->active = 0; }
            }
        }
        return;
    }

    if (phase == PH_FCT_FINISH) {
        // bundle.c:242-275 after the IPM on L0: f = phi + gamma_1 + ... in
        // order (one lane, so it rounds as the reference's loop), g = 1 - B(X)
        if (!p.mask[node]) return;
        const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
        subgradient(d, p, node, s, p.Xt + node * slab, p.g + (size_t)node * d.max_m);
        if (threadIdx.x == 0) {
            const double *gam = p.gamma + (size_t)node * d.max_m;
            double f = p.phi[node];
            for (int c = 0; c < m; ++c) f += gam[c];
            sp->f_test = f;
            sp->status = p.ipm_status[node];
        }
        return;
    }

    if (phase == PH_FINISH) {
        if (threadIdx.x == 0) {
            if (s.active && s.stage == BQB_STAGE_LOOP) { sp->active = 0; sp->bound = node_bound(s, d.certify); }
            p.mask[node] = 0;
        }
        return;
    }
}

// ------------------------------------------------------------------ op_Bt
// Gather lists: for every matrix pair {r,c} the contributions of the cuts
// that contain it, in cut order.  L0(r,c) is then formed by adding those
// contributions to L(r,c) one after the other -- the same sequence of
// additions the reference's scatter performs on that entry.
__global__ __launch_bounds__(CT) void csr_build_kernel(BatchDims d, BatchPtrs p)
{
    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
    int *ptr = p.csr_ptr + (size_t)node * (d.npairs + 1);
    int *fill = p.csr_fill + (size_t)node * d.npairs;
    int *lst = p.csr_cut + (size_t)node * d.max_contrib;
    const TriCut *tri = p.tri + (size_t)node * d.max_tri;
    const KgonCut *pent = p.pent + (size_t)node * d.max_pent;
    const KgonCut *hept = p.hept + (size_t)node * d.max_hept;
    const int np = s.n * (s.n - 1) / 2;
    for (int e = threadIdx.x; e <= np; e += CT) ptr[e] = 0;
    for (int e = threadIdx.x; e < np; e += CT) fill[e] = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < m; c += CT) {
        int pp[7], k;
        if (c < s.NIneq) { pp[0] = tri[c].i; pp[1] = tri[c].j; pp[2] = tri[c].k; k = 3; }
        else if (c < s.NIneq + s.NPentIneq) { for (int a = 0; a < 5; ++a) pp[a] = pent[c - s.NIneq].p[a]; k = 5; }
        else { for (int a = 0; a < 7; ++a) pp[a] = hept[c - s.NIneq - s.NPentIneq].p[a]; k = 7; }
        for (int a = 0; a < k; ++a)
            for (int b = a + 1; b < k; ++b) atomicAdd(&ptr[pair_id(pp[a], pp[b]) + 1], 1);
    }
    __syncthreads();
    // exclusive scan of ptr[1..np] (single wave, chunked)
    if (threadIdx.x < 64) {
        int carry = 0;
        for (int base = 1; base <= np; base += 64) {
            const int i = base + threadIdx.x;
            int v = (i <= np) ? ptr[i] : 0;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(v, o, 64);
                if ((int)threadIdx.x >= o) v += u;
            }
            if (i <= np) ptr[i] = v + carry;
            carry += __shfl(v, 63, 64);
        }
    }
    __syncthreads();
    const int total = ptr[np];
    if (total > d.max_contrib) {
        if (threadIdx.x == 0) { sp->status = ERR_CUTS_FULL; sp->active = 0; p.mask[node] = 0; }
        return;
    }
    for (int c = threadIdx.x; c < m; c += CT) {
        int pp[7], k, type;
        if (c < s.NIneq) { pp[0] = tri[c].i; pp[1] = tri[c].j; pp[2] = tri[c].k; k = 3; type = tri[c].type; }
        else if (c < s.NIneq + s.NPentIneq) { for (int a = 0; a < 5; ++a) pp[a] = pent[c - s.NIneq].p[a]; k = 5; type = pent[c - s.NIneq].type; }
        else { for (int a = 0; a < 7; ++a) pp[a] = hept[c - s.NIneq - s.NPentIneq].p[a]; k = 7; type = hept[c - s.NIneq - s.NPentIneq].type; }
        for (int a = 0; a < k; ++a) {
            const int ea = (k == 3) ? tri_e(type, a) : kgon_e(type, a);
            for (int b = a + 1; b < k; ++b) {
                const int eb = (k == 3) ? tri_e(type, b) : kgon_e(type, b);
                const int pid = pair_id(pp[a], pp[b]);
                const int pos = ptr[pid] + atomicAdd(&fill[pid], 1);
                lst[pos] = (c << 1) | ((ea * eb > 0) ? 0 : 1);   // X += temp (0) or X -= temp (1)
            }
        }
    }
    __syncthreads();
    // order each pair's contributions by cut index
    for (int pid = threadIdx.x; pid < np; pid += CT) {
        const int a0 = ptr[pid], a1 = ptr[pid + 1];
        for (int i = a0 + 1; i < a1; ++i) {
            const int v = lst[i];
            int j = i - 1;
            while (j >= a0 && lst[j] > v) { lst[j + 1] = lst[j]; --j; }
            lst[j + 1] = v;
        }
    }
    if (threadIdx.x == 0) sp->cuts_changed = 0;
}

// L0 = L - B'(gamma_src) for the node (operators.c:140-481): temp =
// 0.5 t (triangle), 0.25 t (pentagonal), (1/6) t (heptagonal).
__global__ __launch_bounds__(CT) void opbt_kernel(BatchDims d, BatchPtrs p, int use_gtest)
{
    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    const NodeState s = p.st[node];
    const size_t slab = (size_t)d.ld * d.ld;
    const double *L = p.L + node * slab;
    double *L0 = p.L0 + node * slab;
    const double *gam = (use_gtest ? p.gtest : p.gamma) + (size_t)node * d.max_m;
    const int *ptr = p.csr_ptr + (size_t)node * (d.npairs + 1);
    const int *lst = p.csr_cut + (size_t)node * d.max_contrib;
    const int n = s.n;
    const int nt = s.NIneq, npt = s.NIneq + s.NPentIneq;
    for (int e = threadIdx.x; e < (int)slab; e += CT) {
        const int r = e % d.ld, c = e / d.ld;
        double v;
        if (r >= n || c >= n) v = 0.0;
        else if (r == c) v = L[e];
        else {
            v = L[e];
            const int pid = pair_id(r, c);
            for (int q = ptr[pid]; q < ptr[pid + 1]; ++q) {
                const int w = lst[q];
                const int cut = w >> 1;
                const double t = gam[cut];
                const double temp = (cut < nt) ? 0.5 * t : (cut < npt) ? 0.25 * t : 1.0 / 6.0 * t;
                v = (w & 1) ? v - temp : v + temp;
            }
        }
        L0[e] = v;
    }
}

}  // namespace bqb

using namespace bqb;

extern "C" hipError_t bqb_launch_control(const BatchDims *d, const BatchPtrs *p, int phase, int arg,
                                         hipStream_t s)
{
    hipLaunchKernelGGL(control_kernel, dim3(d->nodes), dim3(CT), 0, s, *d, *p, phase, arg);
    return hipGetLastError();
}

extern "C" hipError_t bqb_launch_csr_build(const BatchDims *d, const BatchPtrs *p, hipStream_t s)
{
    hipLaunchKernelGGL(csr_build_kernel, dim3(d->nodes), dim3(CT), 0, s, *d, *p);
    return hipGetLastError();
}

extern "C" hipError_t bqb_launch_opbt(const BatchDims *d, const BatchPtrs *p, int use_gtest, hipStream_t s)
{
    hipLaunchKernelGGL(opbt_kernel, dim3(d->nodes), dim3(CT), 0, s, *d, *p, use_gtest);
    return hipGetLastError();
}
