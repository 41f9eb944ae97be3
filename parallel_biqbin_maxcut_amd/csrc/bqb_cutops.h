// bqb_cutops.h -- the cut operator B and its adjoint on device.
//
// A k-gonal cut (k = 3, 5, 7; src/biqbin.h:81-107) is
//   B_c(X) = coef * sum_{a<b} s_ab X(p_a, p_b) <= 1,   s_ab = -e_a e_b,
// coef = 1, 1/2, 1/3, where e is a +-1 vector fixed by the cut type:
// triangle type 1..4 -> e = (1,1,1), (1,1,-1), (1,-1,1), (-1,1,1)
// (cutting_planes.c:20-39); pentagonal/heptagonal type t -> the first t-1
// entries of e are -1 (H = e e', cutting_planes.c:239-244, 431-438).
// Pairs are visited in the order (0,1),(0,2),...,(1,2),... -- the order of
// the terms written out in operators.c -- so sums round identically.
#pragma once
#include "bqb_batch.h"
#include "bqb_round.h"

namespace bqb {

// The node's upper bound f + fixedvalue (bounding.c:177, 339); when the
// context certifies, the certified f (rounded up) takes the place of f.
__device__ __forceinline__ double node_bound(const NodeState &s, int certify)
{
    return certify ? bqb_add_ru(s.f_cert, s.fixedvalue) : s.f + s.fixedvalue;
}

// The node's bundle is full: it leaves SDPbound as a give-up with the bound
// of its current stability center, f + fixedvalue (a valid bound for any
// multipliers gamma >= 0), and the event is flagged.  One thread calls it.
__device__ __forceinline__ void bundle_full_giveup(NodeState *sp, const NodeState &s, int *mask, int certify)
{
    sp->giveup = 1;
    sp->done = 1;
    sp->active = 0;
    sp->bound = node_bound(s, certify);
    sp->flags = s.flags | BQB_FLAG_BUNDLE_FULL;
    *mask = 0;
}

__device__ __forceinline__ int tri_e(int type, int a)
{
    // type 2 flips e[2], type 3 flips e[1], type 4 flips e[0]
    return ((type == 2 && a == 2) || (type == 3 && a == 1) || (type == 4 && a == 0)) ? -1 : 1;
}
__device__ __forceinline__ int kgon_e(int type, int a) { return (a < type - 1) ? -1 : 1; }

// sum_{a<b} s_ab X(p_a, p_b) in operators.c order; X column-major with ld.
__device__ __forceinline__ double cut_sum(const double *X, int ld, const int *p, int k, int type)
{
    double s = 0.0;
    bool first = true;
    for (int a = 0; a < k; ++a) {
        const int ea = (k == 3) ? tri_e(type, a) : kgon_e(type, a);
        for (int b = a + 1; b < k; ++b) {
            const int eb = (k == 3) ? tri_e(type, b) : kgon_e(type, b);
            const double v = X[p[a] + (size_t)p[b] * ld];
            const double tv = (-ea * eb > 0) ? v : -v;
            if (first) { s = tv; first = false; }
            else s = s + tv;
        }
    }
    return s;
}

// B_c(X) of cut c of a node (c indexes [tri | pent | hept]).
__device__ __forceinline__ double cut_B(const double *X, int ld, int c, int ntri, int npent,
                                        const TriCut *tri, const KgonCut *pent, const KgonCut *hept)
{
    if (c < ntri) {
        const TriCut tc = tri[c];
        const int p[3] = {tc.i, tc.j, tc.k};
        return cut_sum(X, ld, p, 3, tc.type);
    }
    if (c < ntri + npent) {
        const KgonCut kc = pent[c - ntri];
        return 0.5 * cut_sum(X, ld, kc.p, 5, kc.type);
    }
    const KgonCut kc = hept[c - ntri - npent];
    return 1.0 / 3.0 * cut_sum(X, ld, kc.p, 7, kc.type);
}

// index of the unordered pair {r, c}, r != c, in the strict upper triangle
__device__ __forceinline__ int pair_id(int r, int c)
{
    if (r > c) { int t = r; r = c; c = t; }
    return c * (c - 1) / 2 + r;
}

// Walks the n x n interior of a column-major slab (leading dimension ld)
// with nt threads: thread t visits flat indices t, t + nt, ... of the n x n
// block (column by column) and off() gives the slab offset r + c ld.  The
// padding outside the interior is never read or written by these walks.
struct InteriorWalk {
    int r, c, dr, dc, n;
    __device__ InteriorWalk(int n_, int t, int nt) : n(n_)
    {
        if (n_ <= 0) { r = dr = dc = 0; c = n_; return; }
        c = t / n_;
        r = t - c * n_;
        dc = nt / n_;
        dr = nt - dc * n_;
    }
    __device__ bool more() const { return c < n; }
    __device__ size_t off(int ld) const { return (size_t)r + (size_t)c * ld; }
    __device__ void next()
    {
        r += dr;
        c += dc;
        if (r >= n) { r -= n; ++c; }
    }
};

}  // namespace bqb
