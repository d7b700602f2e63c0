// BDYGEN (pot/boundary.f:88-908): multipole boundary potential.
//
// The reference loops over every ring (JJ,KK) and, for each of the 383
// boundary points, adds the ring's (l,m) moments either to MASSIN (ring inside
// the point's spherical radius) or, scaled by r^-(2l+1), to MASSOUT
// (boundary.f:556-604): 1.64 G branchy adds at 256x128x256.  The split only
// depends on the ring's spherical radius, so here the rings are sorted by
// radius once per grid; MASSIN of a boundary point is then a prefix sum and
// MASSOUT a suffix sum over the sorted rings.
//
// The ring moments SUMASS, CM(m), SM(m) (boundary.f:526-538) are azimuthal
// Fourier coefficients at the cell-centre angles; the forward transform of the
// interior lanes writes them (fft.hip), so the density is read once per solve.
// Kernels here:
//   k_bdy_chunks   per chunk of <= 64 radius-sorted rings: inside- and
//                  outside-form sums of TERM(l,m) x moment for the l+m-even
//                  terms (boundary.f:541-553 and :585-593)
//   k_bdy_scan     exclusive prefix (inside) / suffix (outside) over chunks;
//                  chunk boundaries include every boundary point's split, so
//                  MASSIN/MASSOUT of a point are one entry each
//   k_bdy_eval     TRMTOT and PHI at the boundary points for all L
//                  (boundary.f:659-830)
//   k_bdy_ghosts   ISYM=2 ghosts (boundary.f:876-893)
// Every reduction has a fixed order that depends on the grid only, so the
// result does not depend on what else runs on the device.
#include "chypot_impl.h"

namespace chy {
namespace {

constexpr int NCOL = 2 * kMaxTerms;  // (term, cos|sin) columns
constexpr int CHUNK_GROUPS = 7;      // ring groups per chunk block (7 x 36 <= 256 threads)

struct BdyArgs {
    double *phi;
    const double *cosm, *sinm;
    const double *mom;            // [nring][kMoments] by ring index
    const int *ring_of_pos;       // radius-sorted position -> ring index
    const double *tin, *tout, *rsph;
    const int *chunk_start;
    double *chunk;
    double *massio;
    const int *jk_cut;
    const double *tb, *rb;
    const int *jpos, *kpos;
    double *dscal;
    int J, K, L, JS, KS, nring, nchunk, jkmax, nt, isym;
    double redge, grav, pi;
    double cnmax, cmmax;          // boundary coefficients of the mode RHS (pot3.f:184-191)
};

// field f's arrays (f: the block's field index)
__device__ __forceinline__ void bind(BdyArgs &a, const Fields &F, int f)
{
    a.phi = F.phi[f];
    a.mom = F.mom[f];
    a.chunk = F.chunk[f];
    a.massio = F.massio[f];
    a.dscal = F.dscal[f];
}

// chunk[0][2t+s][c]: sum over chunk c's rings of TERM_t * moment_s (inside
// form), chunk[1][2t+s][c]: the same with TERM_t r^-(2l+1) (outside form);
// s=0 cosine (SUMASS for m=0), s=1 sine.
// blockIdx.x is the field and blockIdx.y the chunk, so the fields' blocks of
// one chunk run together and share the chunk's rows of the grid's tables
// (TERM, r) in L2 instead of each field streaming the whole tables again.
__global__ __launch_bounds__(256) void k_bdy_chunks(BdyArgs a, const Fields F)
{
    bind(a, F, blockIdx.x);
    constexpr int IT = (kChunkRings + CHUNK_GROUPS - 1) / CHUNK_GROUPS;  // rings per thread
    __shared__ double smom[kChunkRings][kMoments + 1];
    __shared__ double part[CHUNK_GROUPS][kMaxTerms][4];
    const int c = blockIdx.y;
    const int tid = threadIdx.x;
    const int p0 = a.chunk_start[c], p1 = a.chunk_start[c + 1];
    const int t = tid % kMaxTerms, grp = tid / kMaxTerms;
    const bool worker = tid < CHUNK_GROUPS * kMaxTerms;
    // this thread's ring terms first (positions clamped into the chunk) ...
    double ti[IT], to[IT], rs[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int pos = p0 + grp + CHUNK_GROUPS * it;
        const int pc = (worker && pos < p1) ? pos : p0;
        ti[it] = a.tin[(size_t)pc * kMaxTerms + t];
        to[it] = a.tout[(size_t)pc * kMaxTerms + t];
        rs[it] = a.rsph[pc];
    }
    // ... and the moments of the block's rings (ring order in memory, radius
    // order in LDS): positions, then ring indices, then moments, each step's
    // loads all issued before any is used
    constexpr int GPER = (kChunkRings * kMoments + 255) / 256;
    const int ne = (p1 - p0) * kMoments;
    int gr[GPER];
#pragma unroll
    for (int i = 0; i < GPER; ++i) {
        const int v = tid + 256 * i;
        gr[i] = a.ring_of_pos[p0 + (v < ne ? v / kMoments : 0)];
    }
    double gm[GPER];
#pragma unroll
    for (int i = 0; i < GPER; ++i) {
        const int v = tid + 256 * i;
        gm[i] = a.mom[(size_t)gr[i] * kMoments + (v < ne ? v % kMoments : 0)];
    }
#pragma unroll
    for (int i = 0; i < GPER; ++i) {
        const int v = tid + 256 * i;
        if (v < ne) smom[v / kMoments][v % kMoments] = gm[i];
    }
    __syncthreads();
    if (worker) {
        const int m = kTermM[t];
        double ic = 0.0, is = 0.0, oc = 0.0, os = 0.0;
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int r = grp + CHUNK_GROUPS * it;
            // rings beyond REDGE are skipped (boundary.f:379)
            const bool use = t < a.nt && p0 + r < p1 && !(a.redge > 0.0 && rs[it] > a.redge);
            const int rr = use ? r : 0;
            const double mc = smom[rr][m];                       // SUMASS (m=0) or CM(m)
            const double ms = m == 0 ? 0.0 : smom[rr][10 + m];   // SM(m)
            ic = use ? fma(ti[it], mc, ic) : ic;
            is = use ? fma(ti[it], ms, is) : is;
            oc = use ? fma(to[it], mc, oc) : oc;
            os = use ? fma(to[it], ms, os) : os;
        }
        part[grp][t][0] = ic;
        part[grp][t][1] = is;
        part[grp][t][2] = oc;
        part[grp][t][3] = os;
    }
    __syncthreads();
    if (tid < kMaxTerms * 4) {
        const int tt = tid / 4, v = tid % 4;
        double sum = 0.0;
#pragma unroll
        for (int g2 = 0; g2 < CHUNK_GROUPS; ++g2) sum += part[g2][tt][v];
        const int comp = v >> 1, cs = v & 1;
        a.chunk[((size_t)comp * NCOL + 2 * tt + cs) * a.nchunk + c] = sum;
    }
}

// massio[c][0][col] = sum_{c'<c} chunk[c'][0][col]     (c = 0..nchunk)
// massio[c][1][col] = sum_{c'>=c} chunk[c'][1][col]
// One block per (column, component) reads its column (contiguous in the
// [comp][col][chunk] layout) 1024 chunks at a time, every load of a pass
// issued first: four consecutive chunks per thread summed in order, the
// thread totals scanned within each wavefront by shuffles, then the wavefront
// totals in a fixed order.
__global__ __launch_bounds__(256) void k_bdy_scan(BdyArgs a, const Fields F)
{
    bind(a, F, blockIdx.y);
    constexpr int E = 4;
    __shared__ double wsum[4];
    const int col = blockIdx.x % NCOL, comp = blockIdx.x / NCOL;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = a.nchunk;
    const double *src = a.chunk + ((size_t)comp * NCOL + col) * n;
    double carry = 0.0;
    for (int base = 0; base < n; base += 256 * E) {
        double v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = base + E * tid + e;  // position in scan order
            const int c = comp == 0 ? i : n - 1 - i;
            v[e] = src[i < n ? c : 0];
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (base + E * tid + e >= n) v[e] = 0.0;
            if (e > 0) v[e] += v[e - 1];
        }
        double incl = v[E - 1];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        double excl = __shfl_up(incl, 1, 64);
        if (lane == 0) excl = 0.0;
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        double pre = carry;
        for (int x2 = 0; x2 < w; ++x2) pre += wsum[x2];
        pre += excl;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = base + E * tid + e;
            if (i < n) {
                const double s = pre + v[e];
                if (comp == 0) a.massio[((size_t)(i + 1) * 2 + 0) * NCOL + col] = s;  // prefix of c' <= i
                else a.massio[((size_t)(n - 1 - i) * 2 + 1) * NCOL + col] = s;      // suffix from c = n-1-i
            }
        }
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (tid == 0) {
        if (comp == 0) a.massio[col] = 0.0;
        else a.massio[((size_t)n * 2 + 1) * NCOL + col] = 0.0;
        if (comp == 0 && col == 0) a.dscal[0] = carry;  // TMASS = sum of MRING(1,1)
    }
}

__global__ __launch_bounds__(256) void k_bdy_eval(BdyArgs a, const Fields F)
{
    bind(a, F, blockIdx.y);
    __shared__ double t1[kMaxTerms], t2[kMaxTerms];
    const int jk = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < kMaxTerms) {
        double v1 = 0.0, v2 = 0.0;
        if (tid < a.nt) {
            const double rs = a.rb[jk];
            const int l = kTermL[tid];
            double rout = 1.0;
            for (int i = 0; i < l; ++i) rout *= rs;
            const double rin = 1.0 / (rout * rs);  // 1/RSPHER**(L+1)
            const double term = a.tb[(size_t)jk * kMaxTerms + tid];
            const double b = term * rin, h = term * rout;
            const int c = a.isym > 0 ? a.jk_cut[jk] : a.nchunk;
            const double *mi = a.massio + ((size_t)c * 2 + 0) * NCOL;
            const double *mo = a.massio + ((size_t)c * 2 + 1) * NCOL;
            if (a.isym > 0) {
                v1 = b * mi[2 * tid] + h * mo[2 * tid];
                v2 = b * mi[2 * tid + 1] + h * mo[2 * tid + 1];
            } else {  // all mass inside (boundary.f:795-805)
                v1 = b * mi[2 * tid];
                v2 = b * mi[2 * tid + 1];
            }
        }
        t1[tid] = v1;
        t2[tid] = v2;
    }
    __syncthreads();
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t pbase = (size_t)(a.jpos[jk] - 1) + (size_t)a.JS * (a.kpos[jk] - 1);
    for (int l = tid; l < a.L; l += blockDim.x) {
        double cr[10], sr[10];
#pragma unroll
        for (int m = 0; m < 10; ++m) {
            cr[m] = a.cosm[l * 10 + m];
            sr[m] = a.sinm[l * 10 + m];
        }
        // SUM over (l,m) in the reference's order (boundary.f:819-828); terms
        // beyond MAXTRM hold 0 and leave the sum unchanged
        double sum = 0.0;
#pragma unroll
        for (int t = 0; t < kMaxTerms; ++t) {
            if (kTermM[t] == 0) sum += t1[t];
            else sum += 2.0 * (cr[kTermM[t] - 1] * t1[t] + sr[kTermM[t] - 1] * t2[t]);
        }
        a.phi[pbase + plane * l] = -a.grav * sum;
    }
}

// The boundary potential of BDYGEN is a Fourier series in psi with orders
// m <= MAXTRM (boundary.f:814-829); its packed azimuthal spectrum (what POT3's
// forward transform of the boundary lanes yields, pot3.f:92-140) follows in
// closed form.  With C_m = sum_{t: m(t)=m} TRMTOT(1,t), S_m likewise for
// TRMTOT(2,t), and psi_n = (n+1/2) DTHETA:
//   a_0 = -2 G C_0
//   a_m = -2 G ( C_m cos(m DTHETA/2) + S_m sin(m DTHETA/2))     1 <= m <= MAXTRM
//   b_m = -2 G (-C_m sin(m DTHETA/2) + S_m cos(m DTHETA/2))
// and every other coefficient is 0.  Valid for LMAX/2 > MAXTRM.  Writes the
// spectrum of the boundary lane and its COEF entries (pot3.f:105-132).
// TRMTOT(1,t), TRMTOT(2,t) of boundary point jk for term t (boundary.f:782-806)
__device__ __forceinline__ void point_terms(const BdyArgs &a, int jk, int t, double &v1, double &v2)
{
    v1 = 0.0;
    v2 = 0.0;
    if (t >= a.nt) return;
    const double rs = a.rb[jk];
    const double term = a.tb[(size_t)jk * kMaxTerms + t];
    const int c = a.isym > 0 ? a.jk_cut[jk] : a.nchunk;
    const int l = kTermL[t];
    double rout = 1.0;
    for (int i = 0; i < l; ++i) rout *= rs;
    const double rin = 1.0 / (rout * rs);
    const double b = term * rin, h = term * rout;
    const double *mi = a.massio + ((size_t)c * 2 + 0) * NCOL;
    const double *mo = a.massio + ((size_t)c * 2 + 1) * NCOL;
    if (a.isym > 0) {
        v1 = b * mi[2 * t] + h * mo[2 * t];
        v2 = b * mi[2 * t + 1] + h * mo[2 * t + 1];
    } else {  // all mass inside (boundary.f:795-805)
        v1 = b * mi[2 * t];
        v2 = b * mi[2 * t + 1];
    }
}

// C_m, S_m (m = 0..10): the TRMTOT sums grouped by azimuthal order
__device__ __forceinline__ void order_sums(const double *t1, const double *t2, int m, double &c, double &s)
{
    c = 0.0;
    s = 0.0;
#pragma unroll
    for (int t = 0; t < kMaxTerms; ++t)
        if (kTermM[t] == m) { c += t1[t]; s += t2[t]; }
}

// packed coefficient q of the boundary series with sums cmv/smv
__device__ __forceinline__ double packed_coef(const double *cmv, const double *smv, const double2 *twh, int q,
                                              int N, double g2)
{
    if (q == 0) return g2 * cmv[0];
    if (q <= kMaxTrm) return g2 * (cmv[q] * twh[q].x + smv[q] * twh[q].y);
    if (q > N && q - N <= kMaxTrm) {
        const int m = q - N;
        return g2 * (-cmv[m] * twh[m].y + smv[m] * twh[m].x);
    }
    return 0.0;
}

// With `edges`, the block also applies the boundary terms of the mode
// right-hand side (pot3.f:212-224) to the interior cells next to its point:
// a top point J=j+2 <= JMAX takes CNMAX times its own coefficient off
// (J, KMAX); an outer point K=k+2 < KMAX takes CMMAX times its own off
// (JMAX, K); the top point J=JMAX also evaluates the outer point K=KMAX
// (second wavefront) and takes both terms off the corner (JMAX, KMAX), top
// first.  Every interior cell has one writer.  The cells' old values are
// loaded at the start, with the point's tables.
constexpr int SPEC_QP = 4;  // packed coefficients per thread: LMAX <= 1024 over 256 threads
__global__ __launch_bounds__(256) void k_bdy_spec(BdyArgs a, const Fields F, const double2 *twh, int coef_mode,
                                                  int edges)
{
    bind(a, F, blockIdx.y);
    double *spec = static_cast<double *>(F.spec[blockIdx.y]);
    double *coef = coef_mode == CHY_COEF_NONE ? nullptr : F.coef[blockIdx.y];
    __shared__ double t1[2][kMaxTerms], t2[2][kMaxTerms];
    __shared__ double cmv[2][kMaxTrm + 1], smv[2][kMaxTrm + 1];
    const int jk = blockIdx.x;
    const int tid = threadIdx.x;
    const int N = a.L / 2;
    const int j = a.jpos[jk] - 2, k = a.kpos[jk] - 2;
    const int NZ = a.K - 1, M = a.J - 1;
    const bool top = k == NZ;
    const bool corner = edges && top && j == M - 1;
    const size_t JK = (size_t)a.J * a.K, sidx = (size_t)k * a.J + j;
    // the interior cell that takes this point's edge term, if any
    long ycell = -1;
    if (edges) {
        if (top && j <= M - 1) ycell = (long)(NZ - 1) * a.J + j;
        else if (!top && k < NZ - 1) ycell = (long)k * a.J + (M - 1);
    }
    double yv[SPEC_QP];
#pragma unroll
    for (int i = 0; i < SPEC_QP; ++i) {
        const int q = tid + 256 * i;
        const size_t cell = ycell >= 0 ? (size_t)ycell : sidx;
        yv[i] = spec[(size_t)(q < a.L ? q : 0) * JK + cell];
    }
    const int half = tid >> 6, tt = tid & 63;
    if (tt < kMaxTerms && (half == 0 || (half == 1 && corner))) {
        double v1, v2;
        point_terms(a, half == 0 ? jk : a.jkmax - 1, tt, v1, v2);  // second: outer point K=KMAX
        t1[half][tt] = v1;
        t2[half][tt] = v2;
    }
    __syncthreads();
    if (tt <= kMaxTrm && half < 2) order_sums(t1[half], t2[half], tt, cmv[half][tt], smv[half][tt]);
    __syncthreads();
    const double g2 = -2.0 * a.grav;
    const double a0 = g2 * cmv[0][0];
#pragma unroll
    for (int i = 0; i < SPEC_QP; ++i) {
        const int q = tid + 256 * i;
        if (q >= a.L) break;
        const double v = packed_coef(cmv[0], smv[0], twh, q, N, g2);
        spec[(size_t)q * JK + sidx] = v;
        if (ycell >= 0) {
            double y = yv[i];
            if (top) {
                y -= a.cnmax * v;
                if (corner) y -= a.cmmax * packed_coef(cmv[1], smv[1], twh, q, N, g2);
            } else {
                y -= a.cmmax * v;
            }
            spec[(size_t)q * JK + (size_t)ycell] = y;
        }
    }
    if (coef && (coef_mode == CHY_COEF_FULL || (coef_mode == CHY_COEF_K2 && k == 0))) {
        // raw REALTR coefficients are the packed ones times L (SF1 = 1/L)
        const size_t plane = (size_t)a.JS * a.KS;
        const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);
        const double A0 = a0 * a.L, x1 = fabs(A0), pi = a.pi;
        for (int m = 1 + tid; m <= N; m += blockDim.x) {
            double A = 0.0, B = 0.0;
            if (m <= kMaxTrm && m < N) {
                A = a.L * g2 * (cmv[0][m] * twh[m].x + smv[0][m] * twh[m].y);
                B = a.L * g2 * (-cmv[0][m] * twh[m].y + smv[0][m] * twh[m].x);
            }
            double *c1 = coef + base + plane * (size_t)(m - 1);
            double *c2 = c1 + plane * (size_t)N;
            if (m < N) {
                if (x1 != 0.0) *c1 = 2.0 * sqrt(A * A + B * B) / x1;
                double philag;
                if (A > 0.0) philag = atan(B / A);
                else if (A < 0.0) philag = atan(B / A) + pi;
                else if (B > 0.0) philag = 0.5 * pi;
                else if (B < 0.0) philag = -0.5 * pi;
                else philag = 4.0 * pi;
                *c2 = philag * 180.0 / pi;
            } else {
                if (x1 != 0.0) *c1 = fabs(A) / x1;
                *c2 = A0;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_bdy_ghosts(BdyArgs a, const Fields F)
{
    bind(a, F, blockIdx.y);
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

bool bdygen_spectral_ok(const chy_grid *g) { return g->L / 2 > kMaxTrm; }

int launch_bdygen_from_moments_f(const chy_grid *g, const Fields &F, int maxtrm, int isym, double redge,
                                 bool spectral, int coef_mode, hipStream_t s)
{
    int nt = 0;
    while (nt < kMaxTerms && kTermL[nt] <= maxtrm) ++nt;
    BdyArgs a{};
    a.cosm = g->d.cosm; a.sinm = g->d.sinm;
    a.ring_of_pos = g->d.ring_of_pos;
    a.tin = g->d.tin; a.tout = g->d.tout; a.rsph = g->d.rsph;
    a.chunk_start = g->d.chunk_start;
    a.jk_cut = g->d.jk_cut; a.tb = g->d.tb; a.rb = g->d.rb;
    a.jpos = g->d.jpos; a.kpos = g->d.kpos;
    a.J = g->J; a.K = g->K; a.L = g->L; a.JS = g->J + 2; a.KS = g->K + 2;
    a.nring = g->nring; a.nchunk = g->nchunk; a.jkmax = g->jkmax; a.nt = nt; a.isym = isym;
    a.redge = redge; a.grav = g->grav; a.pi = g->pi;
    a.cnmax = g->cnmax; a.cmmax = g->cmmax;
    const int nf = F.nf;
    hipLaunchKernelGGL(k_bdy_chunks, dim3(nf, g->nchunk), dim3(256), 0, s, a, F);
    hipLaunchKernelGGL(k_bdy_scan, dim3(2 * NCOL, nf), dim3(256), 0, s, a, F);
    if (spectral) {
        hipLaunchKernelGGL(k_bdy_spec, dim3(g->jkmax, nf), dim3(256), 0, s, a, F, g->d.twh, coef_mode, 1);
    } else {
        hipLaunchKernelGGL(k_bdy_eval, dim3(g->jkmax, nf), dim3(256), 0, s, a, F);
        hipLaunchKernelGGL(k_bdy_ghosts, dim3((g->L + 255) / 256, nf), dim3(256), 0, s, a, F);
    }
    return (int)hipGetLastError();
}

int launch_bdygen_from_moments(const chy_grid *g, int maxtrm, int isym, double redge, double *phi,
                               bool spectral, double *coef, int coef_mode, hipStream_t s)
{
    const Fields F = one_field(g, nullptr, phi, coef, 0);
    return launch_bdygen_from_moments_f(g, F, maxtrm, isym, redge, spectral, coef_mode, s);
}

}  // namespace chy
