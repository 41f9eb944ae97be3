// bqb_sa.hip -- pentagonal (k=5) and heptagonal (k=7) cut update of a batch
// of nodes (reference src/cutting_planes.c:224-602 and
// src/qap_simulated_annealing.c:12-190).
//
// The Pent_Trials*3 (Hepta_Trials*4) simulated-annealing chains of one
// separation round run one per lane.  Chain c = trial*types + type-1 draws
// its random numbers from the counter stream key(round key, c) in the order
// the reference calls rand(), so a chain is a pure function of X and its key.
// X is staged in LDS as a packed triangle.  The chain results are then merged
// into the list in (trial, type) order, purged, deduplicated and appended
// exactly as the reference does.
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_cutops.h"
#include "bqb_kernels.h"
#include "bqb_rng.h"

namespace bqb {

constexpr int ST = 256;
constexpr int SA_MAX_STAGES = 150;

// The first K entries of a chain's permutation, 16 bits each in two 64-bit
// words (entries 0-3, 4-6): read and written at run-time positions by shifts,
// so they stay in registers.
struct Head {
    uint64_t w0, w1;
    __device__ __forceinline__ int get(int i) const
    {
        return (int)(((i < 4) ? (w0 >> (16 * i)) : (w1 >> (16 * (i - 4)))) & 0xffffull);
    }
    __device__ __forceinline__ void set(int i, int v)
    {
        if (i < 4) w0 = (w0 & ~(0xffffull << (16 * i))) | ((uint64_t)v << (16 * i));
        else w1 = (w1 & ~(0xffffull << (16 * (i - 4)))) | ((uint64_t)v << (16 * (i - 4)));
    }
};

// packed symmetric X: element (a, b) at max*(max+1)/2 + min
__device__ __forceinline__ double xs(const double *Xp, int a, int b)
{
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return Xp[hi * (hi + 1) / 2 + lo];
}

// the same element read from the node's X slab (X too large for LDS)
__device__ __forceinline__ double xg(const double *X, int ld, int a, int b)
{
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return X[lo + (size_t)hi * ld];
}

// A chain's permutation: in LDS (8-bit entries: vertices < n <= n0 <= 256)
// when it fits next to the packed X, otherwise in the node's global scratch.
template <bool LDSP>
struct PermRef {
    unsigned char *s;
    int *g;
    __device__ __forceinline__ int get(int i) const { return LDSP ? (int)s[i] : g[i]; }
    __device__ __forceinline__ void set(int i, int v) const
    {
        if (LDSP) s[i] = (unsigned char)v;
        else g[i] = v;
    }
};

template <int K, bool XL, bool LDSP>
__global__ __launch_bounds__(ST) void sa_kernel(BatchDims d, BatchPtrs p)
{
    // packed triangle of X (XL), then the LDS permutations (LDSP)
    extern __shared__ double s_X[];
    __shared__ double s_t0;
    __shared__ int s_cnt[ST];
    __shared__ int s_qn;
    constexpr int NTYPES = (K == 5) ? 3 : 4;

    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    const BqbParams P = *p.params;
    const int n = s.n;
    const int ld = d.ld;
    const double *X = p.X + (size_t)node * ld * ld;
    const int trials = (K == 5) ? P.Pent_Trials : P.Hepta_Trials;
    const int nchains = trials * NTYPES;
    const int list_cap = (K == 5) ? P.PentIneq : P.HeptaIneq;
    if (nchains > BQB_SA_CHAINS || n < K) {
        // fewer vertices than the cut needs: no chain can run (the reference
        // would index out of range); nothing is separated.
        if (threadIdx.x == 0 && nchains > BQB_SA_CHAINS) { sp->status = ERR_SIZE; sp->active = 0; p.mask[node] = 0; }
        if (threadIdx.x == 0 && K == 5) sp->viol5 = 1e9;
        return;
    }
    auto xa = [&](int a, int b) { return XL ? xs(s_X, a, b) : xg(X, ld, a, b); };
    for (int e = threadIdx.x; XL && e < n * (n + 1) / 2; e += ST) {
        int hi = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while (hi * (hi + 1) / 2 > e) --hi;
        while ((hi + 1) * (hi + 2) / 2 <= e) ++hi;
        const int lo = e - hi * (hi + 1) / 2;
        s_X[e] = X[lo + (size_t)hi * ld];
    }
    // initial temperature (qap_simulated_annealing.c:20-31), summed in the
    // reference's order by one lane so it rounds identically
    if (threadIdx.x == 0) {
        double sumX = n;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) sumX += 2 * fabs(X[j + (size_t)i * ld]);
        double t = K * K;
        t *= sumX;
        t /= (n * (n - 1));
        s_t0 = t;
    }
    __syncthreads();

    const uint64_t rkey = (K == 5) ? BQB_KEY_PENT(s.seed, s.count) : BQB_KEY_HEPT(s.seed, s.count);
    const int c = threadIdx.x;
    if (c < nchains) {
        const int type = c % NTYPES + 1;
        const uint64_t key = bqb_key(rkey, (uint64_t)c);
        uint64_t ctr = 0;
        PermRef<LDSP> perm;
        perm.s = reinterpret_cast<unsigned char *>(s_X + (XL ? (size_t)d.n0 * (d.n0 + 1) / 2 : 0)) + (size_t)c * d.n0;
        perm.g = p.perm + ((size_t)node * BQB_SA_CHAINS + c) * ld;
        for (int i = 0; i < n; ++i) perm.set(i, i);
        for (int i = n - 1; i > 0; --i) {
            const int r = bqb_rmod(bqb_draw(key, ctr++), i);
            const int tmp = perm.get(i);
            perm.set(i, perm.get(r));
            perm.set(r, tmp);
        }
        Head pk = {0ull, 0ull}, pb = {0ull, 0ull};
#pragma unroll
        for (int j = 0; j < K; ++j) pk.set(j, perm.get(j));
        int e[7];
#pragma unroll
        for (int j = 0; j < 7; ++j) e[j] = kgon_e(type, j);
        double t1 = s_t0;
        int m1 = n;
        double sol = K;
#pragma unroll
        for (int i = 0; i < K; ++i)
#pragma unroll
            for (int j = i + 1; j < K; ++j) sol += 2 * (e[j] * e[i]) * xa(pk.get(j), pk.get(i));
        double cost = 1e9;
        int not_done = 1, stages = 0, status = 0;
        const bqb_divm dn = bqb_divm_make(n);
        while (not_done > 0) {
            if (stages++ >= SA_MAX_STAGES) { status = ERR_SA_STAGES; break; }
            not_done = 0;
            for (int it = 0; it < m1; ++it) {
                int i1 = bqb_rmod(bqb_draw(key, ctr++), K);
                int i2 = bqb_rmod_by(bqb_draw(key, ctr++), dn);
                if (i2 < i1) { const int tmp = i1; i1 = i2; i2 = tmp; }
                const int p1 = pk.get(i1);
                const int p2 = (i2 < K) ? pk.get(i2) : perm.get(i2);
                const int ei1 = kgon_e(type, i1);
                double delta = 0.0;
#pragma unroll
                for (int i = 0; i < K; ++i)
                    delta += (double)(e[i] * ei1) * (xa(pk.get(i), p2) - xa(pk.get(i), p1));
                if (i2 < K) {
                    const int ei2 = kgon_e(type, i2);
#pragma unroll
                    for (int i = 0; i < K; ++i)
                        delta += (double)(e[i] * ei2) * (xa(pk.get(i), p1) - xa(pk.get(i), p2));
                }
                delta *= 2.0;
                const double tv1 = 2.0 - 2.0 * xa(p2, p1);
                double tv2 = 1.0;
                if (i2 < K) tv2 += -2 * (kgon_e(type, i1) * kgon_e(type, i2)) + 1.0;
                delta += tv1 * tv2;
                const double sol_temp = sol + delta;
                int accept;
                if (delta > 0) {
                    const double dt1 = delta / t1;
                    if (dt1 > 5) accept = 0;
                    else accept = (bqb_u01(bqb_draw(key, ctr++)) < exp(-dt1)) ? 1 : 0;
                } else
                    accept = 1;
                if (accept) {
                    if (fabs(delta) > 0.001) not_done = 1;
                    pk.set(i1, p2);
                    if (i2 < K) pk.set(i2, p1);
                    else perm.set(i2, p1);
                    sol = sol_temp;
                    if (sol < cost) {
                        cost = sol;
                        pb = pk;
                    }
                }
            }
            t1 *= 0.6;
            m1 = (int)round(m1 * 1.1);
        }
        p.sa_val[(size_t)node * BQB_SA_CHAINS + c] = cost;
#pragma unroll
        for (int j = 0; j < K; ++j) p.sa_perm[((size_t)node * BQB_SA_CHAINS + c) * 7 + j] = pb.get(j);
        p.sa_status[(size_t)node * BQB_SA_CHAINS + c] = status;
    }
    __syncthreads();

    // ---- merge chains into the list in (trial, type) order ---------------
    // (cutting_planes.c:247-304 / 441-498, including the rescan that starts
    // from BIG_NUMBER with a '>' test and so never moves the slot)
    KgonCut *lst = p.kg_lst + (size_t)node * d.kg_list;
    double *lv = p.kg_lval + (size_t)node * d.kg_list;
    if (threadIdx.x == 0) {
        int size = 0, least = 0, st = 0;
        double leastv = -1e9, minall = 1e9;
        for (int q = 0; q < nchains; ++q) {
            const double v = p.sa_val[(size_t)node * BQB_SA_CHAINS + q];
            if (p.sa_status[(size_t)node * BQB_SA_CHAINS + q]) st = p.sa_status[(size_t)node * BQB_SA_CHAINS + q];
            const int type = q % NTYPES + 1;
            minall = (v < minall) ? v : minall;
            if (v < 1.0) {
                int slot = -1;
                if (size < list_cap && size < d.kg_list) {
                    slot = size;
                    if (size == 0 || v > leastv) { leastv = v; least = size; }
                    ++size;
                } else if (v < leastv) {
                    slot = least;
                    leastv = 1e9;
                    for (int l = 0; l < size; ++l)
                        if ((l == slot ? v : lv[l]) > leastv) { leastv = (l == slot ? v : lv[l]); least = l; }
                }
                if (slot >= 0) {
                    lst[slot].type = type;
                    for (int j = 0; j < 7; ++j)
                        lst[slot].p[j] = (j < K) ? p.sa_perm[((size_t)node * BQB_SA_CHAINS + q) * 7 + j] : 0;
                    lv[slot] = v;
                }
            }
        }
        s_cnt[0] = size;
        s_qn = st;
        if (K == 5) sp->viol5 = minall;
    }
    __syncthreads();
    const int size = s_cnt[0];
    if (s_qn) { if (threadIdx.x == 0) { sp->status = s_qn; sp->active = 0; p.mask[node] = 0; } return; }
    __syncthreads();

    // ---- purge (cutting_planes.c:325-350 / 518-543) -----------------------
    KgonCut *cuts = (K == 5) ? p.pent + (size_t)node * d.max_pent : p.hept + (size_t)node * d.max_hept;
    double *cy = (K == 5) ? p.pent_y + (size_t)node * d.max_pent : p.hept_y + (size_t)node * d.max_hept;
    const int ncur = (K == 5) ? s.NPentIneq : s.NHeptaIneq;
    const int off = (K == 5) ? s.off_pent : s.off_hept;
    const int maxc = (K == 5) ? d.max_pent : d.max_hept;
    const double *gam = p.gamma + (size_t)node * d.max_m;
    int kept = 0;
    for (int base = 0; base < ncur; base += ST) {
        const int cc = base + threadIdx.x;
        int keep = 0;
        KgonCut kc;
        double y = 0.0;
        if (cc < ncur) { y = gam[off + cc]; kc = cuts[cc]; keep = !(y < 1e-5); }
        s_cnt[threadIdx.x] = keep;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int i = 0; i < ST; ++i) { const int v = s_cnt[i]; s_cnt[i] = acc; acc += v; }
            s_qn = acc;
        }
        __syncthreads();
        if (keep) { cuts[kept + s_cnt[threadIdx.x]] = kc; cy[kept + s_cnt[threadIdx.x]] = y; }
        kept += s_qn;
        __syncthreads();
    }
    // ---- dedupe against kept cuts, append with y = 0 ----------------------
    int added = 0;
    for (int base = 0; base < size; base += ST) {
        const int l = base + threadIdx.x;
        int fresh = 0;
        KgonCut nc;
        if (l < size) {
            nc = lst[l];
            fresh = 1;
            for (int q = 0; q < kept; ++q) {
                const KgonCut e = cuts[q];
                bool same = e.type == nc.type;
                for (int j = 0; j < K; ++j) same = same && (e.p[j] == nc.p[j]);
                if (same) { fresh = 0; break; }
            }
        }
        s_cnt[threadIdx.x] = fresh;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int i = 0; i < ST; ++i) { const int v = s_cnt[i]; s_cnt[i] = acc; acc += v; }
            s_qn = acc;
        }
        __syncthreads();
        if (fresh) {
            const int pos = kept + added + s_cnt[threadIdx.x];
            if (pos < maxc) { cuts[pos] = nc; cy[pos] = 0.0; }
        }
        added += s_qn;
        __syncthreads();
    }
    const bool trunc = kept + added > maxc;   // more new cuts than the context holds
    if (trunc) added = maxc - kept;
    if (threadIdx.x == 0) {
        if (trunc) sp->flags = s.flags | ((K == 5) ? BQB_FLAG_PENT_TRUNC : BQB_FLAG_HEPT_TRUNC);
        if (K == 5) sp->NPentIneq = kept + added;
        else sp->NHeptaIneq = kept + added;
        sp->cuts_changed = 1;
    }
}

}  // namespace bqb

template <int K>
static void launch_sa(const BatchDims *d, const BatchPtrs *p, bool xl, bool lp, size_t lds, hipStream_t s)
{
    if (xl && lp) hipLaunchKernelGGL((bqb::sa_kernel<K, true, true>), dim3(d->nodes), dim3(bqb::ST), lds, s, *d, *p);
    else if (xl) hipLaunchKernelGGL((bqb::sa_kernel<K, true, false>), dim3(d->nodes), dim3(bqb::ST), lds, s, *d, *p);
    else if (lp) hipLaunchKernelGGL((bqb::sa_kernel<K, false, true>), dim3(d->nodes), dim3(bqb::ST), lds, s, *d, *p);
    else hipLaunchKernelGGL((bqb::sa_kernel<K, false, false>), dim3(d->nodes), dim3(bqb::ST), lds, s, *d, *p);
}

extern "C" hipError_t bqb_launch_sa(const BatchDims *d, const BatchPtrs *p, int k, hipStream_t s)
{
    // sized by n0 (every node has n <= n0): the packed triangle of X, then
    // one byte per vertex and chain
    const size_t lds_x = sizeof(double) * (size_t)d->n0 * (d->n0 + 1) / 2;
    const int chains = (k == 5) ? d->sa_chains5 : d->sa_chains7;
    const size_t lds_p = sizeof(unsigned char) * (size_t)chains * d->n0;
    // X is staged in LDS as a packed triangle while it fits in the 160 KB of
    // a CU (less the kernel's static LDS), else read from the node's slab;
    // the chains' permutations join it in LDS while both fit, otherwise they
    // stay in the node's global scratch
    const size_t cap = 156 * 1024;
    const bool xl = lds_x <= cap;
    const bool lp = (xl ? lds_x : 0) + lds_p <= cap;
    const size_t lds = (xl ? lds_x : 0) + (lp ? lds_p : 0);
    if (k == 5) launch_sa<5>(d, p, xl, lp, lds, s);
    else launch_sa<7>(d, p, xl, lp, lds, s);
    return hipGetLastError();
}
