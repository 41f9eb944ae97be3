// bqb_trisep.hip -- triangle cut update of a batch of nodes
// (reference src/cutting_planes.c:20-214):
//   1. purge: store the current multipliers with the cuts, drop y < 1e-5;
//   2. scan all 4*C(n,3) triangle inequalities of X in the reference's order
//      (type, ii > jj > kk) and keep at most TriIneq violated ones
//      (value > violated_TriIneq) exactly as getViolated_TriangleInequalities
//      does: the first TriIneq fill the list, each later one that beats the
//      current least violated entry replaces it (first minimum in list order);
//   3. append the list entries that are not already kept cuts, with y = 0.
// Evaluation is parallel over the workgroup; the violated candidates of each
// chunk are compacted in scan order into LDS and replayed by one wave, so the
// list (set and order) is the reference's.
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_kernels.h"
#include "bqb_wave.h"

namespace bqb {

constexpr int TT = 256;            // threads
constexpr int PER = 8;             // consecutive triples per thread per chunk
constexpr int CHUNK = TT * PER;
constexpr int TRI_LIST_MAX = BQB_TRI_LIST_MAX;   // LDS list capacity

__device__ __forceinline__ double tri_value(const double *X, int ld, int type, int ii, int jj, int kk)
{
    const double a = X[ii + (size_t)jj * ld], b = X[ii + (size_t)kk * ld], c = X[jj + (size_t)kk * ld];
    switch (type) {
    case 1: return -a - b - c - 1.0;
    case 2: return -a + b + c - 1.0;
    case 3: return a - b + c - 1.0;
    default: return a + b - c - 1.0;
    }
}

__device__ __forceinline__ long long c3(long long n) { return n * (n - 1) * (n - 2) / 6; }

// LM: the list capacity held in LDS -- TRI_LIST_MAX, or 1024 when TriIneq
// is at most that (10 n0 for n0 <= 102), which halves the kernel's LDS so
// that four workgroups fit on a CU.  The same arithmetic either way.
template <int LM>
__global__ __launch_bounds__(TT) void trisep_kernel(BatchDims d, BatchPtrs p)
{
    __shared__ double s_lval[LM];
    __shared__ unsigned s_lidx[LM];      // packed (type-1, ii, jj, kk)
    __shared__ double s_qval[CHUNK];
    __shared__ unsigned s_qidx[CHUNK];
    __shared__ int s_cnt[TT];
    __shared__ double s_red[TT / 64];
    __shared__ int s_size, s_least, s_qn;
    __shared__ double s_leastv;
    __shared__ int s_any[2][TT / 64];   // per-wave "chunk has a violation", double buffered

    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    const BqbParams P = *p.params;
    const int n = s.n;
    const int ld = d.ld;
    const double *X = p.X + (size_t)node * ld * ld;
    TriCut *cuts = p.tri + (size_t)node * d.max_tri;
    double *cy = p.tri_y + (size_t)node * d.max_tri;
    const double *gam = p.gamma + (size_t)node * d.max_m;
    const int cap = P.TriIneq < LM ? P.TriIneq : LM;
    const double thr = P.violated_TriIneq;

    // ---- 1. purge (cutting_planes.c:140-166), chunked stable compaction --
    int kept = 0;
    for (int base = 0; base < s.NIneq; base += TT) {
        const int c = base + threadIdx.x;
        int keep = 0;
        TriCut cc;
        double y = 0.0;
        if (c < s.NIneq) {
            y = gam[c];
            cc = cuts[c];
            keep = !(y < 1e-5);
        }
        s_cnt[threadIdx.x] = keep;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int i = 0; i < TT; ++i) { const int v = s_cnt[i]; s_cnt[i] = acc; acc += v; }
            s_qn = acc;
        }
        __syncthreads();
        if (keep) {
            cuts[kept + s_cnt[threadIdx.x]] = cc;
            cy[kept + s_cnt[threadIdx.x]] = y;
        }
        kept += s_qn;
        __syncthreads();
    }

    // ---- 2. scan ---------------------------------------------------------
    if (threadIdx.x == 0) { s_size = 0; s_least = 0; s_leastv = 1e9; }
    double vmax = -1e9;
    // wave 0, once the list is full: lane b keeps the first minimum (value,
    // position) of list block b (positions 64b .. 64b+63), so a replacement
    // rescans one block and the block minima instead of the whole list
    double bmv = 1e300;
    int bms = 0x7fffffff;
    bool blocks_ready = false;
    const long long per_type = c3(n);
    const long long total = 4 * per_type;
    int parity = 0;
    for (long long q0 = 0; q0 < total; q0 += CHUNK, parity ^= 1) {
        double v[PER];
        unsigned id[PER];
        int nv = 0;
        // decode the first of this thread's PER consecutive triples, then step
        // through (type, ii > jj > kk) in scan order
        const long long qs = q0 + (long long)threadIdx.x * PER;
        int type = 1, ii = 2, jj = 1, kk = 0;
        if (qs < total) {
            type = (int)(qs / per_type) + 1;
            long long r = qs - (long long)(type - 1) * per_type;
            // ii: largest with C(ii,3) <= r
            ii = (int)cbrt(6.0 * (double)r) + 1;
            while (ii > 2 && c3(ii) > r) --ii;
            while (c3(ii + 1) <= r) ++ii;
            r -= c3(ii);
            jj = (int)((1.0 + sqrt(1.0 + 8.0 * (double)r)) * 0.5);
            while (jj > 1 && (long long)jj * (jj - 1) / 2 > r) --jj;
            while ((long long)(jj + 1) * jj / 2 <= r) ++jj;
            kk = (int)(r - (long long)jj * (jj - 1) / 2);
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const long long q = qs + u;
            v[u] = -1e300;
            id[u] = 0;
            if (q < total) {
                const double val = tri_value(X, ld, type, ii, jj, kk);
                vmax = fmax(vmax, val);
                if (val > thr) {
                    v[u] = val;
                    id[u] = ((unsigned)(type - 1) << 30) | ((unsigned)ii << 20) | ((unsigned)jj << 10) | (unsigned)kk;
                    ++nv;
                }
                if (++kk == jj) {
                    kk = 0;
                    if (++jj == ii) {
                        jj = 1;
                        if (++ii == n) { ii = 2; ++type; }
                    }
                }
            }
        }
        // one barrier when no triple of the chunk is violated (the usual case)
        {
            const unsigned long long b = __ballot(nv > 0);
            if ((threadIdx.x & 63) == 0) s_any[parity][threadIdx.x >> 6] = b != 0ull;
            __syncthreads();
            int any = 0;
#pragma unroll
            for (int w = 0; w < TT / 64; ++w) any |= s_any[parity][w];
            if (!any) continue;
        }
        // ordered compaction of the chunk's violated triples
        s_cnt[threadIdx.x] = nv;
        __syncthreads();
        if (threadIdx.x < 64) {
            int carry = 0;
            for (int b = 0; b < TT; b += 64) {
                int x = s_cnt[b + threadIdx.x];
                int incl = x;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const int w = __shfl_up(incl, o, 64);
                    if ((int)threadIdx.x >= o) incl += w;
                }
                s_cnt[b + threadIdx.x] = carry + incl - x;
                carry += __shfl(incl, 63, 64);
            }
            if (threadIdx.x == 0) s_qn = carry;
        }
        __syncthreads();
        {
            int o = s_cnt[threadIdx.x];
#pragma unroll
            for (int u = 0; u < PER; ++u)
                if (v[u] > thr) { s_qval[o] = v[u]; s_qidx[o] = id[u]; ++o; }
        }
        __syncthreads();
        // replay (cutting_planes.c:70-116) by wave 0
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            const int qn = s_qn;
            int size = s_size, least = s_least;
            double leastv = s_leastv;
            int qi = 0;
            // fill phase: slots size.. in order; least = first strict minimum
            while (qi < qn && size < cap) {
                const int take = min(min(64, qn - qi), cap - size);
                double val = 1e300;
                int slot = 0x7fffffff;
                if (lane < take) {
                    val = s_qval[qi + lane];
                    slot = size + lane;
                    s_lval[slot] = val;
                    s_lidx[slot] = s_qidx[qi + lane];
                }
                // first minimum among the new entries
                wave_argmin64(val, slot);
                if (size == 0 || val < leastv) { leastv = val; least = slot; }
                size += take;
                qi += take;
            }
            // replacement phase
            if (qi < qn && !blocks_ready) {
                for (int b = 0; b * 64 < cap; ++b) {
                    const int pos = 64 * b + lane;
                    double bv = (pos < cap) ? s_lval[pos] : 1e300;
                    int bp = (pos < cap) ? pos : 0x7fffffff;
                    wave_argmin64(bv, bp);
                    if (lane == b) { bmv = bv; bms = bp; }
                }
                blocks_ready = true;
            }
            while (qi < qn) {
                // 64 candidates in registers; each replacement re-tests the
                // ones after it against the new least value
                const int m = min(64, qn - qi);
                const double val = (lane < m) ? s_qval[qi + lane] : -1e300;
                const unsigned vid = (lane < m) ? s_qidx[qi + lane] : 0u;
                unsigned long long live = (m == 64) ? ~0ull : ((1ull << m) - 1ull);
                for (;;) {
                    const unsigned long long bal = __ballot(val > leastv) & live;
                    if (!bal) break;
                    const int first = __ffsll((long long)bal) - 1;
                    const double nv = lane_value(val, first);
                    const unsigned nid = (unsigned)lane_value((int)vid, first);
                    if (lane == 0) {
                        s_lval[least] = nv;
                        s_lidx[least] = nid;
                    }
                    __builtin_amdgcn_s_waitcnt(0);
                    __builtin_amdgcn_wave_barrier();
                    // the replaced entry's block, then the first minimum of the
                    // list as the first minimum over the blocks
                    const int b = least >> 6;
                    const int pos = 64 * b + lane;
                    double bv = (pos < cap) ? s_lval[pos] : 1e300;
                    int bp = (pos < cap) ? pos : 0x7fffffff;
                    wave_argmin64(bv, bp);
                    if (lane == b) { bmv = bv; bms = bp; }
                    double mv = bmv;
                    int ms = bms;
                    wave_argmin64(mv, ms);
                    leastv = mv;
                    least = ms;
                    live &= ~((2ull << first) - 1ull);
                }
                qi += m;
            }
            if (lane == 0) { s_size = size; s_least = least; s_leastv = leastv; }
        }
        __syncthreads();
    }
    // maximum violation over all triples
    {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o, 64));
        if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = vmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = s_red[0];
            for (int w = 1; w < TT / 64; ++w) m = fmax(m, s_red[w]);
            s_leastv = m;   // reuse as broadcast slot
        }
        __syncthreads();
    }
    const double maxall = s_leastv;
    const int size = s_size;

    // ---- 3. dedupe against the kept cuts and append (cutting_planes.c:173-208)
    int added = 0;
    for (int base = 0; base < size; base += TT) {
        const int l = base + threadIdx.x;
        int fresh = 0;
        TriCut nc;
        if (l < size) {
            const unsigned w = s_lidx[l];
            nc.type = (int)(w >> 30) + 1;
            nc.i = (int)((w >> 20) & 1023);
            nc.j = (int)((w >> 10) & 1023);
            nc.k = (int)(w & 1023);
            fresh = 1;
            for (int c = 0; c < kept; ++c) {
                const TriCut e = cuts[c];
                if (e.type == nc.type && e.i == nc.i && e.j == nc.j && e.k == nc.k) { fresh = 0; break; }
            }
        }
        s_cnt[threadIdx.x] = fresh;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int i = 0; i < TT; ++i) { const int v = s_cnt[i]; s_cnt[i] = acc; acc += v; }
            s_qn = acc;
        }
        __syncthreads();
        if (fresh) {
            const int pos = kept + added + s_cnt[threadIdx.x];
            if (pos < d.max_tri) { cuts[pos] = nc; cy[pos] = 0.0; }
        }
        added += s_qn;
        __syncthreads();
    }
    const bool trunc = kept + added > d.max_tri;   // more new cuts than the context holds
    if (trunc) added = d.max_tri - kept;
    // the list itself (values in scan-replay order) for inspection
    TriCut *lst = p.tri_lst + (size_t)node * d.tri_list;
    double *lv = p.tri_lval + (size_t)node * d.tri_list;
    for (int l = threadIdx.x; l < size && l < d.tri_list; l += TT) {
        const unsigned w = s_lidx[l];
        lst[l].type = (int)(w >> 30) + 1;
        lst[l].i = (int)((w >> 20) & 1023);
        lst[l].j = (int)((w >> 10) & 1023);
        lst[l].k = (int)(w & 1023);
        lv[l] = s_lval[l];
    }
    if (threadIdx.x == 0) {
        sp->off_pent = s.NIneq;
        sp->off_hept = s.NIneq + s.NPentIneq;
        sp->NIneq = kept + added;
        if (trunc) sp->flags = s.flags | BQB_FLAG_TRI_TRUNC;
        sp->viol3 = maxall;
        sp->cuts_changed = 1;
    }
}

}  // namespace bqb

// tri_ineq: the TriIneq parameter of the launch's nodes (the list capacity)
extern "C" hipError_t bqb_launch_trisep(const BatchDims *d, const BatchPtrs *p, int tri_ineq, hipStream_t s)
{
    if (tri_ineq <= 1024)
        hipLaunchKernelGGL(bqb::trisep_kernel<1024>, dim3(d->nodes), dim3(bqb::TT), 0, s, *d, *p);
    else
        hipLaunchKernelGGL(bqb::trisep_kernel<bqb::TRI_LIST_MAX>, dim3(d->nodes), dim3(bqb::TT), 0, s, *d, *p);
    return hipGetLastError();
}
