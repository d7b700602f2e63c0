// BDYGEN (pot/boundary.f:88-908): multipole boundary potential.
//
// The reference loops over every ring (JJ,KK) and, for each of the 383
// boundary points, adds the ring's 66 (l,m) moments either to MASSIN (ring
// inside the point's spherical radius) or, scaled by r^-(2l+1), to MASSOUT
// (boundary.f:556-604): 1.64 G branchy adds at 256x128x256.  The split only
// depends on the ring's spherical radius, so here the rings are sorted by
// radius once per grid; then MASSIN of a boundary point is a prefix sum and
// MASSOUT a suffix sum over the sorted rings.  Kernels:
//   k_bdy_moments  ring moments SUMASS, CM(m), SM(m) over L (boundary.f:526-538)
//   k_bdy_chunks   per chunk of <= 64 radius-sorted rings: inside- and
//                  outside-form sums of TERM(l,m) x moment for the 36 l+m-even
//                  terms (boundary.f:541-553 and :585-593)
//   k_bdy_scan     exclusive prefix (inside) / suffix (outside) over chunks;
//                  chunk boundaries contain every boundary point's split, so
//                  MASSIN/MASSOUT of point jk are one entry each
//   k_bdy_eval     TRMTOT and PHI at the boundary points for all L
//                  (boundary.f:659-830)
//   k_bdy_ghosts   ISYM=2 ghosts (boundary.f:876-893)
// Every reduction has a fixed order that depends on the grid only, so the
// result does not depend on what else runs on the device.
#include "chypot_impl.h"

namespace chy {
namespace {

constexpr int NCOL = 2 * kMaxTerms;  // (term, cos|sin) columns

struct BdyArgs {
    const double *rho;
    double *phi;
    const double *dv;
    const double *cosm, *sinm;
    double *mom;
    const int *ring_of_pos;
    const double *tin, *tout, *rsph;
    const int *chunk_start;
    double *chunk;
    double *massio;
    const int *jk_cut;
    const double *tb, *rb;
    const int *jpos, *kpos;
    double *dscal;
    int J, K, L, JS, KS, nring, nchunk, jkmax, nt, isym;
    double redge, grav;
    int term_l[kMaxTerms], term_m[kMaxTerms];  // (l,m) of each term, l-major
};

__global__ __launch_bounds__(256) void k_bdy_moments(BdyArgs a)
{
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir >= a.nring) return;
    const int jj = 2 + ir % (a.J - 1), kk = 2 + ir / (a.J - 1);
    const size_t plane = (size_t)a.JS * a.KS;
    const double *rp = a.rho + (size_t)(jj - 1) + (size_t)a.JS * (kk - 1);
    const double dv = a.dv[ir];
    double s = 0.0, cmv[10], smv[10];
#pragma unroll
    for (int m = 0; m < 10; ++m) { cmv[m] = 0.0; smv[m] = 0.0; }
    for (int l = 0; l < a.L; ++l) {
        const double h = rp[plane * l] * dv;
        s += h;
#pragma unroll
        for (int m = 0; m < 10; ++m) {
            cmv[m] = fma(a.cosm[l * 10 + m], h, cmv[m]);
            smv[m] = fma(a.sinm[l * 10 + m], h, smv[m]);
        }
    }
    double *o = a.mom + (size_t)ir * kMoments;
    o[0] = s;
#pragma unroll
    for (int m = 0; m < 10; ++m) {
        o[1 + m] = cmv[m];
        o[11 + m] = smv[m];
    }
}

// chunk[c][0][col]: sum TERM*moment (inside form), chunk[c][1][col]: outside form
__global__ __launch_bounds__(128) void k_bdy_chunks(BdyArgs a)
{
    const int c = blockIdx.x;
    const int col = threadIdx.x;
    if (col >= NCOL) return;
    const int t = col >> 1, cs = col & 1;
    double ai = 0.0, ao = 0.0;
    if (t < a.nt) {
        const int m = a.term_m[t];
        const int midx = m == 0 ? (cs ? -1 : 0) : (cs ? 10 + m : m);
        const int p0 = a.chunk_start[c], p1 = a.chunk_start[c + 1];
        for (int pos = p0; pos < p1; ++pos) {
            if (a.redge > 0.0 && a.rsph[pos] > a.redge) continue;  // boundary.f:379
            const int ir = a.ring_of_pos[pos];
            const double mv = midx < 0 ? 0.0 : a.mom[(size_t)ir * kMoments + midx];
            ai = fma(a.tin[(size_t)pos * kMaxTerms + t], mv, ai);
            ao = fma(a.tout[(size_t)pos * kMaxTerms + t], mv, ao);
        }
    }
    a.chunk[((size_t)c * 2 + 0) * NCOL + col] = ai;
    a.chunk[((size_t)c * 2 + 1) * NCOL + col] = ao;
}

// massio[c][0][col] = sum_{c'<c} chunk[c'][0][col]     (c = 0..nchunk)
// massio[c][1][col] = sum_{c'>=c} chunk[c'][1][col]
// one block per (column, component); Hillis-Steele scans of 256 in LDS
__global__ __launch_bounds__(256) void k_bdy_scan(BdyArgs a)
{
    __shared__ double buf[256];
    const int col = blockIdx.x % NCOL, comp = blockIdx.x / NCOL;
    const int tid = threadIdx.x;
    const int n = a.nchunk;
    double carry = 0.0;
    for (int base = 0; base < n; base += 256) {
        const int i = base + tid;  // position in scan order
        double v = 0.0;
        if (i < n) {
            const int c = comp == 0 ? i : n - 1 - i;
            v = a.chunk[((size_t)c * 2 + comp) * NCOL + col];
        }
        buf[tid] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const double add = tid >= off ? buf[tid - off] : 0.0;
            __syncthreads();
            buf[tid] += add;
            __syncthreads();
        }
        const double incl = carry + buf[tid];
        const double excl = incl - v;  // not used for sums; recomputed below
        (void)excl;
        if (i < n) {
            if (comp == 0) {
                // exclusive prefix up to c=i is the inclusive prefix up to i-1
                a.massio[((size_t)(i + 1) * 2 + 0) * NCOL + col] = incl;
            } else {
                const int c = n - 1 - i;
                a.massio[((size_t)c * 2 + 1) * NCOL + col] = incl;
            }
        }
        carry += buf[255];
        __syncthreads();
    }
    if (tid == 0) {
        if (comp == 0) a.massio[(size_t)0 * 2 * NCOL + col] = 0.0;
        else a.massio[((size_t)n * 2 + 1) * NCOL + col] = 0.0;
        if (comp == 0 && col == 0) a.dscal[0] = carry;  // TMASS = sum of MRING(1,1)
    }
}

__global__ __launch_bounds__(256) void k_bdy_eval(BdyArgs a)
{
    __shared__ double t1[kMaxTerms], t2[kMaxTerms];
    __shared__ int tm[kMaxTerms];
    const int jk = blockIdx.x;
    const int tid = threadIdx.x;
    const double rs = a.rb[jk];
    if (tid < a.nt) {
        const int l = a.term_l[tid];
        double rout = 1.0;
        for (int i = 0; i < l; ++i) rout *= rs;
        const double rin = 1.0 / (rout * rs);  // 1/RSPHER**(L+1)
        const double term = a.tb[(size_t)jk * kMaxTerms + tid];
        const double b = term * rin, h = term * rout;
        const int c = a.isym > 0 ? a.jk_cut[jk] : a.nchunk;
        const double *mi = a.massio + ((size_t)c * 2 + 0) * NCOL;
        const double *mo = a.massio + ((size_t)c * 2 + 1) * NCOL;
        if (a.isym > 0) {
            t1[tid] = b * mi[2 * tid] + h * mo[2 * tid];
            t2[tid] = b * mi[2 * tid + 1] + h * mo[2 * tid + 1];
        } else {
            t1[tid] = b * mi[2 * tid];
            t2[tid] = b * mi[2 * tid + 1];
        }
        tm[tid] = a.term_m[tid];
    }
    __syncthreads();
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t pbase = (size_t)(a.jpos[jk] - 1) + (size_t)a.JS * (a.kpos[jk] - 1);
    for (int l = tid; l < a.L; l += blockDim.x) {
        double sum = 0.0;
        for (int t = 0; t < a.nt; ++t) {
            const int m = tm[t];
            if (m == 0) sum += t1[t];
            else sum += 2.0 * (a.cosm[l * 10 + m - 1] * t1[t] + a.sinm[l * 10 + m - 1] * t2[t]);
        }
        a.phi[pbase + plane * l] = -a.grav * sum;
    }
}

__global__ __launch_bounds__(256) void k_bdy_ghosts(BdyArgs a)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= a.L) return;
    const size_t plane = (size_t)a.JS * a.KS;
    const int n = a.L / 2;
    // PHI(JMAX+1,1,LL) = PHI(JMAX+1,2,LL)
    a.phi[(size_t)a.J + plane * l] = a.phi[(size_t)a.J + (size_t)a.JS + plane * l];
    // PHI(1,KMAX+1,LL) = PHI(2,KMAX+1,LL+-LMAX/2)
    const int lp = l < n ? l + n : l - n;
    const size_t rowk = (size_t)a.JS * a.K;
    a.phi[rowk + plane * l] = a.phi[rowk + 1 + plane * lp];
}

}  // namespace

int launch_bdygen(const chy_grid *g, int maxtrm, int isym, double redge, const double *rho,
                  double *phi, hipStream_t s)
{
    int n = 0;
    const Term *tt = terms_table(&n);
    int nt = 0;
    while (nt < n && tt[nt].l <= maxtrm) ++nt;
    BdyArgs a{};
    a.rho = rho; a.phi = phi;
    a.dv = g->d.dv; a.cosm = g->d.cosm; a.sinm = g->d.sinm;
    a.mom = g->d.mom; a.ring_of_pos = g->d.ring_of_pos;
    a.tin = g->d.tin; a.tout = g->d.tout; a.rsph = g->d.rsph;
    a.chunk_start = g->d.chunk_start; a.chunk = g->d.chunk; a.massio = g->d.massio;
    a.jk_cut = g->d.jk_cut; a.tb = g->d.tb; a.rb = g->d.rb;
    a.jpos = g->d.jpos; a.kpos = g->d.kpos; a.dscal = g->d.dscal;
    a.J = g->J; a.K = g->K; a.L = g->L; a.JS = g->J + 2; a.KS = g->K + 2;
    a.nring = g->nring; a.nchunk = g->nchunk; a.jkmax = g->jkmax; a.nt = nt; a.isym = isym;
    a.redge = redge; a.grav = g->grav;
    for (int i = 0; i < kMaxTerms; ++i) { a.term_l[i] = tt[i].l; a.term_m[i] = tt[i].m; }
    hipLaunchKernelGGL(k_bdy_moments, dim3((g->nring + 255) / 256), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_bdy_chunks, dim3(g->nchunk), dim3(128), 0, s, a);
    hipLaunchKernelGGL(k_bdy_scan, dim3(2 * NCOL), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_bdy_eval, dim3(g->jkmax), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_bdy_ghosts, dim3((g->L + 255) / 256), dim3(256), 0, s, a);
    return (int)hipGetLastError();
}

}  // namespace chy
