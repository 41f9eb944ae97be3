// bqb_cert.hip -- certified upper bounds of the dual function.
//
// For multipliers gamma >= 0 of valid cuts and any y, weak duality gives
//   max { <L, X> : diag X = e, B(X) <= 1, X psd }
//     <= sum(gamma) + e'y + n max(0, -lambda_min(Z)),  Z = Diag(y) - L + B'(gamma)
// (for X feasible, <L, X> = <Diag y, X> - <Z, X> + <B'gamma, X>
//  <= e'y - lambda_min(Z) tr X + gamma'B(X)).  The interior-point solve keeps
// Z positive definite in floating point; the certificate turns that into a
// rigorous bound with arithmetic rounded toward +inf / -inf (bqb_round.h):
//   1. Z~ = fl(Z) entry by entry, with an enclosure [lo, hi] of the exact
//      entry: E1 = Z - Z~, |E1_ij| <= e1_ij.
//   2. A Cholesky factor R of Z^ = Z~ + delta I (delta = 0 first, then small
//      shifts) in round-to-nearest; d >= max_i (Z^_ii - Z~_ii).
//   3. E2 = R R' - Z^ enclosed entry by entry with directed rounding.
//   4. Z = R R' - E2 - (Z^ - Z~) + E1 >= -(d + r) I with r >= the largest row
//      sum of |E1| + |E2| (>= the spectral norm of the symmetric error).
// The certified value is RU(sum gamma + e'y + n (d + r)); if no shift gives
// a factor, Gershgorin's bound on the enclosure is used instead.  The
// reference's bounding.c:109, 180, 197 prune on f + fixedvalue; with
// certification the prune tests use this value instead of f.
//
// One workgroup of 256 threads per node; Z~ and R live in two of the node's
// IPM work slabs (the IPM is not running while a certificate is computed).
#include <hip/hip_runtime.h>

#include "bqb_batch.h"
#include "bqb_cutops.h"
#include "bqb_kernels.h"
#include "bqb_round.h"

namespace bqb {

constexpr int CE = 256;   // threads of a certificate workgroup

// Contributions of the cuts to the off-diagonal entry (i, j) of L0 = L - B'gamma:
// the op_Bt gather list of the pair, in cut order (operators.c:140-481).
struct CutSource {
    const int *ptr, *lst;      // gather offsets / entries (cut << 1 | minus)
    const double *gam;         // multipliers of the point certified
    int ntri, ntp;             // triangle count, triangle + pentagonal count
};

// Enclosure of L0_ij: value in round-to-nearest (the op_Bt order) and
// [lo, hi] around the exact value.
__device__ __forceinline__ void l0_entry(const CutSource *cs, double L, int i, int j, double &rn, double &lo,
                                         double &hi)
{
    rn = lo = hi = L;
    if (!cs) return;
    const int pid = pair_id(i, j);
    for (int q = cs->ptr[pid]; q < cs->ptr[pid + 1]; ++q) {
        const int w = cs->lst[q];
        const int cut = w >> 1;
        // the bundle's multipliers can carry rounding-level negatives
        // (gamma + t (eta - G lambda) cancels where eta > 0); weak duality
        // needs gamma >= 0, so the certificate is that of max(gamma, 0)
        const double t = fmax(cs->gam[cut], 0.0);
        double t_rn, t_lo, t_hi;
        if (cut < cs->ntri) { t_rn = t_lo = t_hi = 0.5 * t; }
        else if (cut < cs->ntp) { t_rn = t_lo = t_hi = 0.25 * t; }
        else { t_rn = 1.0 / 6.0 * t; t_lo = bqb_div_rd(t, 6.0); t_hi = bqb_div_ru(t, 6.0); }
        if (w & 1) {
            rn = rn - t_rn;
            lo = bqb_sub_rd(lo, t_hi);
            hi = bqb_sub_ru(hi, t_lo);
        } else {
            rn = rn + t_rn;
            lo = bqb_add_rd(lo, t_lo);
            hi = bqb_add_ru(hi, t_hi);
        }
    }
}

__device__ __forceinline__ double red_ru(double v, double *red)
{
    // upward sum over the workgroup: every step rounds up, so the result is
    // an upper bound of the exact sum whatever the order
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = bqb_add_ru(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < CE / 64; ++w) s = bqb_add_ru(s, red[w]);
    __syncthreads();
    return s;
}

__device__ __forceinline__ double red_max(double v, double *red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < CE / 64; ++w) s = fmax(s, red[w]);
    __syncthreads();
    return s;
}

// The certificate of one node: returns RU(sum_gamma_ru + e'y + n (d + r)).
// L (ld) the node's objective, y its dual vector, cs the cuts (NULL: none).
// Zt: an ld x ld work slab; R: the factor's storage with leading dimension
// ldr (a second work slab, or LDS for small n).  rs: LDS of 2 ld doubles.
__device__ double certify(int n, int ld, const double *L, const double *y, const CutSource *cs,
                          double sum_gamma_ru, double *Zt, double *R, int ldr, double *rs, double *red)
{
    const int tid = threadIdx.x;
    // 1. Z~ and the enclosure widths, entry by entry over all threads (the
    // gather lists make each entry a chain of dependent loads); the widths
    // are parked in R (not used before step 2) and summed by rows in column
    // order
    for (int e = tid; e < n * n; e += CE) {
        const int i = e % n, j = e / n;
        const double l = L[i + (size_t)j * ld];
        double zt, e1;
        if (i == j) {
            zt = y[i] - l;
            e1 = fmax(bqb_sub_ru(bqb_sub_ru(y[i], l), zt), bqb_sub_ru(zt, bqb_sub_rd(y[i], l)));
        } else {
            double rn, lo, hi;
            l0_entry(cs, l, i, j, rn, lo, hi);
            zt = -rn;                                       // Z_ij = -L0_ij
            e1 = fmax(bqb_sub_ru(hi, rn), bqb_sub_ru(rn, lo));
        }
        Zt[i + (size_t)j * ld] = zt;
        R[i + (size_t)j * ldr] = e1;
    }
    __syncthreads();
    for (int i = tid; i < n; i += CE) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc = bqb_add_ru(acc, R[i + (size_t)j * ldr]);
        rs[i] = acc;
    }
    __syncthreads();
    double zmax = 0.0;
    for (int i = tid; i < n; i += CE) zmax = fmax(zmax, fabs(Zt[i + (size_t)i * ld]));
    zmax = red_max(zmax, red);

    // 2. Cholesky of Z~ + delta I, delta = 0, then 1e-12, 1e-10, 1e-8 of max |Z~_ii|
    // (right-looking; the trailing update of step k is spread over a 16 x 16
    // thread grid, every entry updated in k order as by a row loop)
    double dshift = 0.0, delta = 0.0;
    bool ok = false;
    const int gi = tid % 16, gj = tid / 16;
    for (int att = 0; att < 4 && !ok; ++att) {
        delta = (att == 0) ? 0.0 : zmax * ((att == 1) ? 1e-12 : (att == 2) ? 1e-10 : 1e-8);
        double dsh = 0.0;
        for (int e = tid; e < n * n; e += CE) {
            const int i = e % n, j = e / n;
            if (j > i) continue;
            double v = Zt[i + (size_t)j * ld];
            if (i == j) {
                const double vd = v + delta;
                dsh = fmax(dsh, bqb_sub_ru(vd, v));
                v = vd;
            }
            R[i + (size_t)j * ldr] = v;
        }
        dshift = red_max(dsh, red);      // (its barriers also publish R)
        ok = true;
        for (int k = 0; k < n; ++k) {
            // every thread reads the pivot (the same value), so the test
            // and the square root need no broadcast
            const double dk = R[k + (size_t)k * ldr];
            if (!(dk > 0.0)) { ok = false; break; }
            const double rk = sqrt(dk);
            for (int i = k + 1 + tid; i < n; i += CE) R[i + (size_t)k * ldr] = R[i + (size_t)k * ldr] / rk;
            __syncthreads();
            if (tid == 0) R[k + (size_t)k * ldr] = rk;     // read again only after the loop
            for (int i = k + 1 + gi; i < n; i += 16) {
                const double rik = R[i + (size_t)k * ldr];
                for (int j = k + 1 + gj; j <= i; j += 16)
                    R[i + (size_t)j * ldr] = fma(-rik, R[j + (size_t)k * ldr], R[i + (size_t)j * ldr]);
            }
            __syncthreads();
        }
        __syncthreads();
    }

    double sy = 0.0;
    for (int i = tid; i < n; i += CE) sy = bqb_add_ru(sy, y[i]);
    sy = red_ru(sy, red);

    double lam;   // lambda_min(Z) >= -lam
    if (ok) {
        // 3. |R R' - Z^| entry by entry with directed rounding: every pair
        // (a, b), a >= b, once (flattened over the threads for balance); the
        // off-diagonal values are parked in the unused upper triangle of R,
        // the diagonal ones in rs2, then summed by rows
        const int np = n * (n + 1) / 2;
        for (int e = tid; e < np; e += CE) {
            int a = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
            while (a * (a + 1) / 2 > e) --a;
            while ((a + 1) * (a + 2) / 2 <= e) ++a;
            const int b = e - a * (a + 1) / 2;
            double hi = 0.0, lo = 0.0;
            for (int k = 0; k <= b; ++k) {
                const double ra = R[a + (size_t)k * ldr], rb = R[b + (size_t)k * ldr];
                hi = bqb_add_ru(hi, bqb_mul_ru(ra, rb));
                lo = bqb_add_rd(lo, bqb_mul_rd(ra, rb));
            }
            double tgt = Zt[a + (size_t)b * ld];
            if (a == b) tgt = tgt + delta;                  // the entry that was factored
            const double e2 = fmax(bqb_sub_ru(hi, tgt), bqb_sub_ru(tgt, lo));
            if (a == b) rs[ld + a] = e2;
            else R[b + (size_t)a * ldr] = e2;
        }
        __syncthreads();
        double rmax = 0.0;
        for (int i = tid; i < n; i += CE) {
            double acc = bqb_add_ru(rs[i], rs[ld + i]);
            for (int j = 0; j < i; ++j) acc = bqb_add_ru(acc, R[j + (size_t)i * ldr]);
            for (int j = i + 1; j < n; ++j) acc = bqb_add_ru(acc, R[i + (size_t)j * ldr]);
            rmax = fmax(rmax, acc);
        }
        rmax = red_max(rmax, red);
        lam = bqb_add_ru(dshift, rmax);
    } else {
        // Gershgorin on the enclosure: lambda_min >= min_i (Z_ii - sum_j |Z_ij|)
        double gmin = INFINITY;
        for (int i = tid; i < n; i += CE) {
            double off = rs[i];                             // widths of row i (diagonal included)
            for (int j = 0; j < n; ++j)
                if (j != i) off = bqb_add_ru(off, fabs(Zt[i + (size_t)j * ld]));
            gmin = fmin(gmin, bqb_sub_rd(Zt[i + (size_t)i * ld], off));
        }
        gmin = -red_max(-gmin, red);
        lam = gmin < 0.0 ? -gmin : 0.0;
    }
    return bqb_add_ru(bqb_add_ru(sum_gamma_ru, sy), bqb_mul_ru((double)n, lam));
}

// The factor of the certificate lives in LDS when the slabs are at most NL
// wide (NL = 64 or 128: 32 or 128 KB), else in the node's second work slab.
template <int NL>
struct CertR {
    __device__ __forceinline__ static double *ptr(double *slab, double *lds) { return NL ? lds : slab; }
};

// ---- the IPM entry point: certificates of e'y for Z = Diag(y) - L ----
template <int NL>
__global__ __launch_bounds__(CE) void cert_ipm_kernel(const int *nvec, const double *L, const double *y,
                                                      long ystride, double *work, int ld, int wslabs,
                                                      const int *status, double *out)
{
    __shared__ double rs[512], red[CE / 64];
    __shared__ double s_R[NL ? NL * NL : 1];
    const int node = blockIdx.x;
    const int n = nvec[node];
    if (status[node] || n < 1 || n > ld || n > 256) { if (threadIdx.x == 0) out[node] = INFINITY; return; }
    const size_t slab = (size_t)ld * ld;
    double *Zt = work + (size_t)node * wslabs * slab;
    const double v = certify(n, ld, L + node * slab, y + node * ystride, nullptr, 0.0, Zt,
                             CertR<NL>::ptr(Zt + slab, s_R), NL ? n : ld, rs, red);
    if (threadIdx.x == 0) out[node] = v;
}

// ---- the node bound: certificate of the point just evaluated ----
// mode 0: basic bound (no cuts; phi of the IPM on L, y in ytest);
// mode 1: the center of a node whose center moved (serious steps, or a
// rebuild that made warm-started multipliers the center) since its last
// certificate, at the end of its round (gamma the center, y of the center's
// evaluation saved in ycen; the cut lists and gather lists are those of the
// evaluations that moved it), the value the round-end prune test reads --
// the same value a certificate after every move would leave there;
// mode 2: right after a rebuild that made the warm-started center gamma the
// new f (y of the IPM on L - B'gamma; the launch trains leave such a center
// to mode 1).
// The certified value replaces NodeState.f_cert.
template <int NL>
__global__ __launch_bounds__(CE) void cert_node_kernel(BatchDims d, BatchPtrs p, int mode)
{
    __shared__ double rs[512], red[CE / 64];
    __shared__ double s_R[NL ? NL * NL : 1];
    const int node = blockIdx.x;
    if (node >= d.nodes || !p.mask[node]) return;
    NodeState *sp = p.st + node;
    const NodeState s = *sp;
    if (!s.active || s.status) return;
    if (mode == 1 && !s.cert_stale) return;
    if (mode == 2 && (s.stage != BQB_STAGE_LOOP || s.rphase != 0)) return;
    const int n = s.n;
    if (n < 1 || n > d.ld) return;
    const size_t slab = (size_t)d.ld * d.ld;
    double *Zt = p.ipm_work + (size_t)node * d.work_slabs * slab;
    const double *y = (mode == 1 ? p.ycen : p.ytest) + (size_t)node * d.ld;
    CutSource cs;
    const CutSource *csp = nullptr;
    double sg = 0.0;
    if (mode >= 1) {
        const int m = s.NIneq + s.NPentIneq + s.NHeptaIneq;
        const double *gam = p.gamma + (size_t)node * d.max_m;
        cs.ptr = p.csr_ptr + (size_t)node * (d.npairs + 1);
        cs.lst = p.csr_cut + (size_t)node * d.max_contrib;
        cs.gam = gam;
        cs.ntri = s.NIneq;
        cs.ntp = s.NIneq + s.NPentIneq;
        csp = &cs;
        double part = 0.0;
        for (int c = threadIdx.x; c < m; c += CE) part = bqb_add_ru(part, fmax(gam[c], 0.0));
        sg = red_ru(part, red);
    }
    const double v = certify(n, d.ld, p.L + node * slab, y, csp, sg, Zt, CertR<NL>::ptr(Zt + slab, s_R),
                             NL ? n : d.ld, rs, red);
    if (threadIdx.x == 0) { sp->f_cert = v; sp->cert_stale = 0; }
}

}  // namespace bqb

using namespace bqb;

extern "C" hipError_t bqb_launch_cert_ipm(const int *nvec, const double *L, const double *y, long ystride,
                                          double *work, int ld, int wslabs, const int *status, double *out,
                                          int nodes, hipStream_t s)
{
    if (ld <= 64)
        hipLaunchKernelGGL(cert_ipm_kernel<64>, dim3(nodes), dim3(CE), 0, s, nvec, L, y, ystride, work, ld, wslabs,
                           status, out);
    else if (ld <= 128)
        hipLaunchKernelGGL(cert_ipm_kernel<128>, dim3(nodes), dim3(CE), 0, s, nvec, L, y, ystride, work, ld, wslabs,
                           status, out);
    else
        hipLaunchKernelGGL(cert_ipm_kernel<0>, dim3(nodes), dim3(CE), 0, s, nvec, L, y, ystride, work, ld, wslabs,
                           status, out);
    return hipGetLastError();
}

extern "C" hipError_t bqb_launch_cert_node(const BatchDims *d, const BatchPtrs *p, int mode, hipStream_t s)
{
    if (d->ld <= 64)
        hipLaunchKernelGGL(cert_node_kernel<64>, dim3(d->nodes), dim3(CE), 0, s, *d, *p, mode);
    else if (d->ld <= 128)
        hipLaunchKernelGGL(cert_node_kernel<128>, dim3(d->nodes), dim3(CE), 0, s, *d, *p, mode);
    else
        hipLaunchKernelGGL(cert_node_kernel<0>, dim3(d->nodes), dim3(CE), 0, s, *d, *p, mode);
    return hipGetLastError();
}
