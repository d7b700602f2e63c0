// Batched azimuthal real FFT of the potential solve.
//
// Forward: for every lane (J,K), J=2..JMAX+1, K=2..KMAX+1, the LMAX real
// samples PHI(J,K,1..LMAX) (RHO on the interior, the BDYGEN boundary values on
// J=JMAX+1 and K=KMAX+1, pot/pot3.f:71-81) are transformed as FFT(+1) followed
// by REALTR(+1) (pot/fft.f:2-70, sign e^{+i theta}) and scaled by
// SF1=0.5/N (pot/pot3.f:60,135-140):
//     a_m = (2/L) sum_n x_n cos(2 pi m n / L),  m = 0..N
//     b_m = (2/L) sum_n x_n sin(2 pi m n / L),  m = 1..N-1
// The COEF diagnostics of pot/pot3.f:105-132 are formed from the same
// coefficients.  Inverse (pot/pot3.f:256-283): the real Fourier series
//     x_n = a_0/2 + sum_{m=1}^{N-1} (a_m cos + b_m sin)(2 pi m n/L) + a_N/2 (-1)^n,
// written with POT3's axis and midplane ghost cells (pot/pot3.f:287-296).
//
// Layout.  Lanes are consecutive J (the contiguous index of PHI), so every
// load and store of a wave touches contiguous J runs.  The spectrum goes to a
// mode-major array spec[q][k][j] (q = packed coefficient index of the
// reference: q<=N holds a_q, q>N holds b_{q-N}); j runs over J=2..JMAX+1 and k
// over K=2..KMAX+1.
//
// Method.  The L=2N real samples are packed as N complex values
// z_n = x_{2n} + i x_{2n+1}; the N-point complex DFT is split N = N1*N2: N2
// threads per lane each do an N1-point DFT in registers, exchange through
// LDS (layout [element][lane], lane fastest: conflict-free), then N2-point
// DFTs in registers, then the real-split post-pass.  All twiddles are tables of
// exact angles.
#include "chypot_impl.h"

namespace chy {
namespace {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
constexpr int bitrevc(int i, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; ++b)
        if (i & (1 << b)) r |= 1 << (bits - 1 - b);
    return r;
}

// v[k] <- sum_n v[n] exp(SIGN 2 pi i n k / P), P a power of two <= 64, radix-2
// decimation in time, fully unrolled (every index is a compile-time constant).
template <int P, int SIGN>
__device__ __forceinline__ void dft_reg(double2 (&v)[P], const double2 *__restrict__ tw64)
{
    if constexpr (P > 1) {
        constexpr int B = ilog2c(P);
        double2 t[P];
#pragma unroll
        for (int i = 0; i < P; ++i) t[bitrevc(i, B)] = v[i];
#pragma unroll
        for (int len = 2; len <= P; len <<= 1) {
#pragma unroll
            for (int i = 0; i < P; i += len) {
#pragma unroll
                for (int jj = 0; jj < len / 2; ++jj) {
                    double2 w = tw64[jj * (64 / len)];
                    if (SIGN < 0) w.y = -w.y;
                    double2 a = t[i + jj];
                    double2 b = jj == 0 ? t[i + jj + len / 2] : cmul(t[i + jj + len / 2], w);
                    t[i + jj] = cadd(a, b);
                    t[i + jj + len / 2] = csub(a, b);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < P; ++i) v[i] = t[i];
    }
}

// Workgroups are dealt to the 8 XCDs round-robin by block index, and each
// XCD has its own L2.  Neighbouring lane chunks share the cache lines at their
// ends (rows of PHI are not 128-byte aligned), so chunk c goes to the block
// that gives every XCD one contiguous run of chunks: those shared lines are
// then fetched (or written back) once.  The grid is rounded up to a multiple
// of 8; blocks past the last chunk leave at once.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_chunk(int nchunks)
{
    const int per = (nchunks + kXcds - 1) / kXcds;
    return (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
}

// the radices of a mixed-radix transform, first pass first
constexpr int kMaxPasses = 12;
struct RadixPlan {
    int n;
    int radix[kMaxPasses];
};

// the power-of-two part of N in passes of 16 (and one of 2, 4 or 8), then
// each odd prime factor; false if N has a prime factor above 23
bool radix_plan(int N, RadixPlan &rp)
{
    rp.n = 0;
    if (N < 1) return false;
    int n = N, p2 = 1;
    while (n % 2 == 0) {
        p2 *= 2;
        n /= 2;
    }
    for (; p2 >= 16; p2 /= 16) rp.radix[rp.n++] = 16;
    if (p2 > 1) rp.radix[rp.n++] = p2;
    for (int p : {3, 5, 7, 11, 13, 17, 19, 23})
        for (; n % p == 0; n /= p) {
            if (rp.n == kMaxPasses) return false;
            rp.radix[rp.n++] = p;
        }
    return n == 1;
}

struct FftArgs {
    const void *rho;      // RHO (forward input on the interior; TI elements)
    const double *phi;    // PHI (forward input on the boundary lanes)
    double *phi_out;      // PHI (inverse output)
    void *spec;           // [L][K][J] of TS, or null (forward: no spectrum written)
    double *coef;         // COEF(JMAX+2,KMAX+2,LMAX/2,2) or null
    double *mom;          // ring moments [nring][21] of interior lanes, or null
    const double *dv;     // DV per ring (boundary.f:516)
    const double2 *twL;   // exp(2 pi i e/L)
    const double2 *tw64;  // exp(2 pi i e/64)
    const double2 *twh;   // exp(pi i m/L), m = 0..10
    int J, K, L, JS, KS;
    int lanes;            // LANES_ALL / LANES_INTERIOR / LANES_BOUNDARY
    int nlanes;
    int coef_mode;
    double sf1, pi;
    double ys;            // factor on the interior lanes' spectrum (4 pi G)
    int ghosts;           // inverse: also write POT3's axis/midplane ghost cells
    double *zsamp;        // inverse: [KMAX][kZaxSamples][L] copy of PHI(J=2..6,K,L) for ZAXPHI, or null
};

// lane index -> (j,k), j = J-2 (0..JMAX-1), k = K-2 (0..KMAX-1)
__device__ __forceinline__ void lane_jk(const FftArgs &a, long lg, int &j, int &k)
{
    if (a.lanes == LANES_ALL) {
        j = (int)(lg % a.J);
        k = (int)(lg / a.J);
    } else if (a.lanes == LANES_INTERIOR) {
        j = (int)(lg % (a.J - 1));
        k = (int)(lg / (a.J - 1));
    } else if (lg < a.J) {  // top row K=KMAX+1, J=2..JMAX+1
        j = (int)lg;
        k = a.K - 1;
    } else {                // outer column J=JMAX+1, K=2..KMAX
        j = a.J - 1;
        k = (int)(lg - a.J);
    }
}

// field f's arrays (blockIdx.y); a pointer left null by the launcher stays null
__device__ __forceinline__ void bind(FftArgs &a, const Fields &F, int f)
{
    if (a.rho) a.rho = F.rho[f];
    if (a.phi) a.phi = F.phi[f];
    if (a.phi_out) a.phi_out = F.phi[f];
    if (a.spec) a.spec = F.spec[f];
    if (a.mom) a.mom = F.mom[f];
    if (a.zsamp) a.zsamp = F.zsamp[f];
    a.coef = a.coef_mode == CHY_COEF_NONE ? nullptr : F.coef[f];
}

// Phase 3 of the forward transform, shared by both kernel families: the real
// split X_m = E_m + e^{2 pi i m/L} O_m of the complex result Z in lds
// ([element][lane], lane fastest), then the ring moments, the spectrum and
// the COEF diagnostics of lane (j,k).  NS threads serve each lane (sub =
// 0..NS-1).
template <typename TS>
__device__ __forceinline__ void forward_outputs(const FftArgs &a, const double2 *lds, const double2 *tw, int N,
                                                int NL, int NS, int lane, int sub, int j, int k, bool interior,
                                                size_t base, size_t plane)
{
    // A_m = 2 Re X_m, B_m = 2 Im X_m
    auto xm = [&](int m) -> double2 {
        double2 zm = lds[(m % N) * NL + lane];
        double2 zc = lds[((N - m) % N) * NL + lane];
        zc.y = -zc.y;
        double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
        double2 d = csub(zm, zc);
        double2 o = make_double2(0.5 * d.y, -0.5 * d.x);
        return cadd(e, cmul(tw[m], o));
    };
    const size_t JK = (size_t)a.J * a.K;
    const size_t sidx = (size_t)k * a.J + j;
    if (a.mom && j <= a.J - 2 && k <= a.K - 2) {
        // ring moments of BDYGEN (boundary.f:526-538) at the cell-centre angles
        // psi_n = (n+1/2) DTHETA:  SUMASS = DV X_0,  CM(m) + i SM(m) = DV e^{i pi m/L} X_m
        const int ir = j + (a.J - 1) * k;
        const double dv = a.dv[ir];
        for (int mm = sub; mm <= kMaxTrm; mm += NS) {
            const int r = mm % a.L;
            double2 X = r <= N ? xm(r) : xm(a.L - r);
            if (r > N) X.y = -X.y;
            const double2 w = a.twh[mm];
            const double re = X.x * w.x - X.y * w.y, im = X.y * w.x + X.x * w.y;
            double *o = a.mom + (size_t)ir * kMoments;  // the lanes' records are contiguous
            if (mm == 0) {
                o[0] = dv * re;
            } else {
                o[mm] = dv * re;
                o[10 + mm] = dv * im;
            }
        }
    }
    const bool want_coef = a.coef && (a.coef_mode == CHY_COEF_FULL || (a.coef_mode == CHY_COEF_K2 && k == 0));
    if (!a.spec && !want_coef) return;
    TS *spec = static_cast<TS *>(a.spec);
    // interior lanes feed the mode right-hand side Y = 4 pi G PHI~ (pot3.f:212)
    const double ys = interior ? a.ys : 1.0;
    double A0 = 0.0;
    if (want_coef) A0 = 2.0 * xm(0).x;
    for (int m = sub; m <= N; m += NS) {
        const double2 X = xm(m);
        const double A = 2.0 * X.x, B = 2.0 * X.y;
        if (spec) {
            spec[(size_t)m * JK + sidx] = (TS)(A * a.sf1 * ys);
            if (m >= 1 && m <= N - 1) spec[(size_t)(N + m) * JK + sidx] = (TS)(B * a.sf1 * ys);
        }
        if (want_coef && m >= 1) {
            // COEF(J,K,m,1) amplitude, COEF(J,K,m,2) phase (pot3.f:105-132)
            const double x1 = fabs(A0);
            double *c1 = a.coef + base + plane * (size_t)(m - 1);
            double *c2 = c1 + plane * (size_t)N;
            if (m < N) {
                if (x1 != 0.0) *c1 = 2.0 * sqrt(A * A + B * B) / x1;
                double philag;
                if (A > 0.0) philag = atan(B / A);
                else if (A < 0.0) philag = atan(B / A) + a.pi;
                else if (B > 0.0) philag = 0.5 * a.pi;
                else if (B < 0.0) philag = -0.5 * a.pi;
                else philag = 4.0 * a.pi;
                *c2 = philag * 180.0 / a.pi;
            } else {
                if (x1 != 0.0) *c1 = fabs(A) / x1;
                *c2 = A0;
            }
        }
    }
}

// Phase 0b of the inverse transform, in place in lds ([element][lane]):
//   Z_k = (C_k + conj C_{N-k}) + i e^{-2 pi i k/L} (C_k - conj C_{N-k}),
// with C_0 = a_0/2 and C_N = a_N/2 packed in slot 0.  The thread that reads
// the pair (k, N-k) also writes it.
__device__ __forceinline__ void inverse_presplit(double2 *lds, const double2 *tw, int N, int NL, int NS, int lane,
                                                 int sub)
{
    for (int kk = sub; kk <= N / 2; kk += NS) {
        if (kk == 0) {
            const double2 c = lds[lane];
            lds[lane] = make_double2(c.x + c.y, c.x - c.y);
        } else {
            const int kn = N - kk;
            const double2 ck = lds[kk * NL + lane], cn = lds[kn * NL + lane];
            const double2 U = make_double2(ck.x + cn.x, ck.y - cn.y);
            const double2 V = make_double2(ck.x - cn.x, ck.y + cn.y);
            double2 w = tw[kk];
            w.y = -w.y;
            const double2 wv = cmul(w, V);
            lds[kk * NL + lane] = make_double2(U.x - wv.y, U.y + wv.x);
            if (kn != kk) lds[kn * NL + lane] = make_double2(U.x + wv.y, -U.y + wv.x);
        }
    }
}

// PHI(J,K,2n+1) = v0 and PHI(J,K,2n+2) = v1 of lane (j,k) (J = j+2,
// K = k+2), with POT3's ghost cells from the same values (pot3.f:287-296):
// the J=2 lanes (axis) also write the axis ghost PHI(1,K,L)=PHI(2,K,L+-N),
// the K=2 lanes (mid) the midplane row PHI(J,1,L)=PHI(J,2,L), the (2,2) lane
// both and PHI(1,1,L); with samples, the lanes J=2..6 also keep ZAXPHI's
// copy, contiguous in L.
__device__ __forceinline__ void store_pair(const FftArgs &a, int n, double v0, double v1, int j, int k, size_t base,
                                           size_t plane, bool axis, bool mid, bool samples, int N, int L)
{
    double *p0 = a.phi_out + base + plane * (size_t)(2 * n);
    p0[0] = v0;
    p0[plane] = v1;
    if (samples && a.zsamp && j < kZaxSamples) {
        double *zp = a.zsamp + ((size_t)k * kZaxSamples + j) * L + 2 * n;
        zp[0] = v0;
        zp[1] = v1;
    }
    if (mid) {
        p0[-a.JS] = v0;
        p0[plane - a.JS] = v1;
    }
    if (axis) {
        const size_t l0 = (size_t)((2 * n + N) % L), l1 = (size_t)((2 * n + 1 + N) % L);
        a.phi_out[base - 1 + plane * l0] = v0;
        a.phi_out[base - 1 + plane * l1] = v1;
        if (mid) {
            a.phi_out[base - 1 - a.JS + plane * l0] = v0;
            a.phi_out[base - 1 - a.JS + plane * l1] = v1;
        }
    }
}

// TI: element type of the forward input on the interior lanes; TS: storage
// type of the spectrum.  The transform itself is computed in binary64.
template <int N1, int N2, int NL, typename TI, typename TS>
__global__ __launch_bounds__(NL *N2) void k_fft_forward(FftArgs a, const Fields F)
{
    bind(a, F, blockIdx.y);
    constexpr int N = N1 * N2;
    constexpr int L = 2 * N;
    __shared__ double2 lds[N * NL];
    __shared__ double2 tw[L];  // exp(2 pi i e/L), staged from a.twL
    const int tid = threadIdx.x;
    const int lane = tid % NL;
    const int sub = tid / NL;
    const int chunk = xcd_chunk((a.nlanes + NL - 1) / NL);
    if (chunk * NL >= a.nlanes) return;  // whole block
    const long lg = (long)chunk * NL + lane;
    const bool active = lg < a.nlanes;
    int j = 0, k = 0;
    lane_jk(a, active ? lg : 0, j, k);  // inactive lanes read lane 0's data and store nothing
    // interior lanes J<=JMAX, K<=KMAX carry the density (pot3.f:71-81)
    const bool interior = j <= a.J - 2 && k <= a.K - 2;
    const TI *rho = static_cast<const TI *>(a.rho);
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);

    // phase 1: N1-point DFTs over z_{N2*n1+n2}, then twiddle W_N^{n2 k1}
    double2 v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int n = N2 * n1 + sub;
        const size_t i0 = base + plane * (size_t)(2 * n), i1 = i0 + plane;
        v[n1] = interior ? make_double2((double)rho[i0], (double)rho[i1]) : make_double2(a.phi[i0], a.phi[i1]);
    }
    for (int e = tid; e < L; e += N2 * NL) tw[e] = a.twL[e];
    __syncthreads();
    dft_reg<N1, +1>(v, a.tw64);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        if (k1 > 0 && sub > 0) v[k1] = cmul(v[k1], tw[(2 * sub * k1) % L]);
        lds[(k1 * N2 + sub) * NL + lane] = v[k1];
    }
    __syncthreads();

    // phase 2: N2-point DFTs; Z[k1 + N1 k2]
    constexpr int R2 = N1 / N2;
    double2 u[R2][N2];
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int k1 = sub + N2 * r;
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) u[r][n2] = lds[(k1 * N2 + n2) * NL + lane];
        dft_reg<N2, +1>(u[r], a.tw64);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int k1 = sub + N2 * r;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) lds[(k1 + N1 * k2) * NL + lane] = u[r][k2];
    }
    __syncthreads();
    if (!active) return;

    forward_outputs<TS>(a, lds, tw, N, NL, N2, lane, sub, j, k, interior, base, plane);
}

// TS: storage type of the spectrum.
template <int N1, int N2, int NL, typename TS>
__global__ __launch_bounds__(NL *N2) void k_fft_inverse(FftArgs a, const Fields F)
{
    bind(a, F, blockIdx.y);
    constexpr int N = N1 * N2;
    constexpr int L = 2 * N;
    constexpr int NM = N / N2 + 1;  // m = sub + N2 t, t < NM, covers 0..N
    __shared__ double2 lds[N * NL];
    __shared__ double2 tw[L];  // exp(2 pi i e/L), staged from a.twL
    const int tid = threadIdx.x;
    const int lane = tid % NL;
    const int sub = tid / NL;
    const int chunk = xcd_chunk((a.nlanes + NL - 1) / NL);
    if (chunk * NL >= a.nlanes) return;  // whole block
    const long lg = (long)chunk * NL + lane;
    const bool active = lg < a.nlanes;
    int j = 0, k = 0;
    lane_jk(a, active ? lg : 0, j, k);  // inactive lanes read lane 0's data and store nothing
    const size_t JK = (size_t)a.J * a.K;
    const size_t sidx = (size_t)k * a.J + j;

    // phase 0: C_0 = a_0/2, C_N = a_N/2 (both real, packed in slot 0),
    //          C_m = (a_m + i b_m)/2, m = 1..N-1.  Every load is issued
    //          before any is used.
    const TS *spec = static_cast<const TS *>(a.spec);
    double av[NM], bv[NM];
#pragma unroll
    for (int t = 0; t < NM; ++t) {
        const int m = sub + N2 * t;
        const int ma = m <= N ? m : N;
        const int mb = (m >= 1 && m <= N - 1) ? N + m : 0;
        av[t] = (double)spec[(size_t)ma * JK + sidx];
        bv[t] = (double)spec[(size_t)mb * JK + sidx];
    }
    for (int e = tid; e < L; e += N2 * NL) tw[e] = a.twL[e];
#pragma unroll
    for (int t = 0; t < NM; ++t) {
        const int m = sub + N2 * t;
        if (m == 0) lds[lane].x = 0.5 * av[t];
        else if (m == N) lds[lane].y = 0.5 * av[t];
        else if (m < N) lds[m * NL + lane] = make_double2(0.5 * av[t], 0.5 * bv[t]);
    }
    __syncthreads();
    inverse_presplit(lds, tw, N, NL, N2, lane, sub);
    __syncthreads();

    // phase 1: z_n = sum_k Z_k e^{-2 pi i k n/N}, split as the forward pass
    double2 v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = lds[(N2 * n1 + sub) * NL + lane];
    dft_reg<N1, -1>(v, a.tw64);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        if (sub > 0) {
            double2 w = tw[(2 * sub * k1) % L];
            w.y = -w.y;
            v[k1] = cmul(v[k1], w);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) lds[(k1 * N2 + sub) * NL + lane] = v[k1];
    __syncthreads();

    constexpr int R2 = N1 / N2;
    double2 u[R2][N2];
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int k1 = sub + N2 * r;
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) u[r][n2] = lds[(k1 * N2 + n2) * NL + lane];
        dft_reg<N2, -1>(u[r], a.tw64);
    }
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);
    const bool axis = a.ghosts && j == 0, mid = a.ghosts && k == 0;
    if (!active) return;
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int k1 = sub + N2 * r;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
            store_pair(a, k1 + N1 * k2, u[r][k2].x, u[r][k2].y, j, k, base, plane, axis, mid, true, N, L);
    }
}

// ---------------------------------------------------------------------------
// Mixed-radix transforms: N = LMAX/2 not a power of two, with prime factors up
// to 23 (the sizes pot/fft.f:111-126 takes besides the powers of two).  A
// lane's N complex values sit in LDS ([element][lane], lane fastest) and go
// through one self-sorting (Stockham) pass per factor, from one buffer to the
// other, so the result comes out in natural order.  Pass s of radix P, with
// Ns the product of the radices before it, takes for every butterfly
// b = 0..N/P-1 the inputs b + r N/P (r < P), turns input r by
// exp(+-2 pi i r (b mod Ns)/(Ns P)), transforms the P values in registers
// (powers of two as above, odd primes as direct DFTs folded on the pairs
// (r, P-r)) and writes output r to (b - b mod Ns) P + b mod Ns + r Ns.  Every
// twiddle is an entry of the exact-angle table exp(2 pi i e/L).

// y_q = sum_r v[r] exp(SIGN 2 pi i r q / P), P an odd prime, written to
// out[q * ostride] as it is formed; tw holds exp(2 pi i e/L) and P divides L.
// v is folded in place into the sums and differences of the pairs (r, P-r).
template <int P, int SIGN>
__device__ __forceinline__ void dft_odd_store(double2 (&v)[P], const double2 *tw, int L, double2 *out, int ostride)
{
    constexpr int H = P / 2;
    const int st = L / P;
    double2 y0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
        const double2 a = v[r], b = v[P - r];
        v[r] = cadd(a, b);
        v[P - r] = csub(a, b);
        y0 = cadd(y0, v[r]);
    }
    out[0] = y0;
    // one output pair at a time: unrolled, the compiler would keep all
    // (P/2)^2 twiddles of the larger primes in registers
#pragma unroll 1
    for (int q = 1; q <= H; ++q) {
        // y_q = v_0 + sum_r (sum_r cos t_rq + i SIGN dif_r sin t_rq), t_rq = 2 pi r q/P;
        // y_{P-q} the same with -SIGN
        double cx = v[0].x, cy = v[0].y, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            const double2 w = tw[((r * q) % P) * st];
            cx += v[r].x * w.x;
            cy += v[r].y * w.x;
            sx += v[P - r].x * w.y;
            sy += v[P - r].y * w.y;
        }
        if (SIGN < 0) {
            sx = -sx;
            sy = -sy;
        }
        out[q * ostride] = make_double2(cx - sy, cy + sx);
        out[(P - q) * ostride] = make_double2(cx + sy, cy - sx);
    }
}

// one Stockham pass of radix P from src to dst (threads sub = 0..NS-1 of
// each lane share the N/P butterflies)
template <int P, int SIGN>
__device__ __forceinline__ void stockham_pass(const double2 *src, double2 *dst, int N, int Ns, int NL, int NS,
                                              int lane, int sub, const double2 *tw, const double2 *tw64, int L)
{
    const int nb = N / P;
    const int tstep = 2 * (N / (Ns * P));  // table step of exp(2 pi i/(Ns P)); r kk tstep < L
    for (int b = sub; b < nb; b += NS) {
        const int kk = b % Ns;
        double2 v[P];
#pragma unroll
        for (int r = 0; r < P; ++r) v[r] = src[(b + r * nb) * NL + lane];
        if (kk > 0) {
#pragma unroll
            for (int r = 1; r < P; ++r) {
                double2 w = tw[r * kk * tstep];
                if (SIGN < 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        const int d = (b - kk) * P + kk;
        if constexpr ((P & (P - 1)) == 0) {
            dft_reg<P, SIGN>(v, tw64);
#pragma unroll
            for (int r = 0; r < P; ++r) dst[(d + r * Ns) * NL + lane] = v[r];
        } else {
            dft_odd_store<P, SIGN>(v, tw, L, dst + d * NL + lane, Ns * NL);
        }
    }
}

// the passes of plan rp, src -> dst -> src ...; returns the buffer holding
// the transform
template <int SIGN>
__device__ __forceinline__ double2 *stockham(const RadixPlan &rp, double2 *src, double2 *dst, int N, int NL, int NS,
                                             int lane, int sub, const double2 *tw, const double2 *tw64, int L)
{
    int Ns = 1;
    for (int s = 0; s < rp.n; ++s) {
        const int P = rp.radix[s];
        switch (P) {
        case 2: stockham_pass<2, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 3: stockham_pass<3, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 4: stockham_pass<4, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 5: stockham_pass<5, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 7: stockham_pass<7, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 8: stockham_pass<8, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 11: stockham_pass<11, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 13: stockham_pass<13, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 16: stockham_pass<16, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 17: stockham_pass<17, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        case 19: stockham_pass<19, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        default: stockham_pass<23, SIGN>(src, dst, N, Ns, NL, NS, lane, sub, tw, tw64, L); break;
        }
        __syncthreads();
        double2 *t = src;
        src = dst;
        dst = t;
        Ns *= P;
    }
    return src;
}

// NL lanes per block, NS threads per lane, N <= NMAX (two buffers of
// NMAX*NL values and the table of L = 2N twiddles in static LDS)
template <int NL, int NS, int NMAX>
__global__ __launch_bounds__(NL *NS) void k_fft_forward_mr(FftArgs a, const Fields F, RadixPlan rp)
{
    bind(a, F, blockIdx.y);
    __shared__ double2 buf[2][NMAX * NL];
    __shared__ double2 tw[2 * NMAX];
    const int L = a.L, N = L / 2;
    const int tid = threadIdx.x;
    const int lane = tid % NL;
    const int sub = tid / NL;
    const int chunk = xcd_chunk((a.nlanes + NL - 1) / NL);
    if (chunk * NL >= a.nlanes) return;  // whole block
    const long lg = (long)chunk * NL + lane;
    const bool active = lg < a.nlanes;
    int j = 0, k = 0;
    lane_jk(a, active ? lg : 0, j, k);  // inactive lanes read lane 0's data and store nothing
    const bool interior = j <= a.J - 2 && k <= a.K - 2;
    const double *rho = static_cast<const double *>(a.rho);
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);
    // z_n = x_{2n} + i x_{2n+1}
    for (int n = sub; n < N; n += NS) {
        const size_t i0 = base + plane * (size_t)(2 * n), i1 = i0 + plane;
        buf[0][n * NL + lane] = interior ? make_double2(rho[i0], rho[i1]) : make_double2(a.phi[i0], a.phi[i1]);
    }
    for (int e = tid; e < L; e += NL * NS) tw[e] = a.twL[e];
    __syncthreads();
    const double2 *z = stockham<+1>(rp, buf[0], buf[1], N, NL, NS, lane, sub, tw, a.tw64, L);
    if (!active) return;
    forward_outputs<double>(a, z, tw, N, NL, NS, lane, sub, j, k, interior, base, plane);
}

template <int NL, int NS, int NMAX>
__global__ __launch_bounds__(NL *NS) void k_fft_inverse_mr(FftArgs a, const Fields F, RadixPlan rp)
{
    bind(a, F, blockIdx.y);
    __shared__ double2 buf[2][NMAX * NL];
    __shared__ double2 tw[2 * NMAX];
    const int L = a.L, N = L / 2;
    const int tid = threadIdx.x;
    const int lane = tid % NL;
    const int sub = tid / NL;
    const int chunk = xcd_chunk((a.nlanes + NL - 1) / NL);
    if (chunk * NL >= a.nlanes) return;  // whole block
    const long lg = (long)chunk * NL + lane;
    const bool active = lg < a.nlanes;
    int j = 0, k = 0;
    lane_jk(a, active ? lg : 0, j, k);  // inactive lanes read lane 0's data and store nothing
    const size_t JK = (size_t)a.J * a.K;
    const size_t sidx = (size_t)k * a.J + j;
    // C_0 = a_0/2 and C_N = a_N/2 packed in slot 0, C_m = (a_m + i b_m)/2
    const double *spec = static_cast<const double *>(a.spec);
    for (int m = sub; m <= N; m += NS) {
        const double am = spec[(size_t)m * JK + sidx];
        if (m == 0) buf[0][lane].x = 0.5 * am;
        else if (m == N) buf[0][lane].y = 0.5 * am;
        else buf[0][m * NL + lane] = make_double2(0.5 * am, 0.5 * spec[(size_t)(N + m) * JK + sidx]);
    }
    for (int e = tid; e < L; e += NL * NS) tw[e] = a.twL[e];
    __syncthreads();
    inverse_presplit(buf[0], tw, N, NL, NS, lane, sub);
    __syncthreads();
    // z_n = sum_k Z_k e^{-2 pi i k n/N}, x_{2n} = Re z_n, x_{2n+1} = Im z_n
    const double2 *z = stockham<-1>(rp, buf[0], buf[1], N, NL, NS, lane, sub, tw, a.tw64, L);
    if (!active) return;
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);
    const bool axis = a.ghosts && j == 0, mid = a.ghosts && k == 0;
    for (int n = sub; n < N; n += NS) {
        const double2 v = z[n * NL + lane];
        store_pair(a, n, v.x, v.y, j, k, base, plane, axis, mid, true, N, L);
    }
}

// transform variants (the refinement works in the spectrum and uses the
// binary64 transforms)
enum { FWD64 = 0, INV64 = 2 };

template <int N1, int N2, int NL, int KIND>
int launch_one(const FftArgs &a, const Fields &F, hipStream_t s)
{
    const int chunks = (a.nlanes + NL - 1) / NL;
    const int blocks = (chunks + kXcds - 1) / kXcds * kXcds;
    const dim3 grid(blocks, F.nf), block(NL * N2);
    if (KIND == FWD64) hipLaunchKernelGGL((k_fft_forward<N1, N2, NL, double, double>), grid, block, 0, s, a, F);
    if (KIND == INV64) hipLaunchKernelGGL((k_fft_inverse<N1, N2, NL, double>), grid, block, 0, s, a, F);
    return (int)hipGetLastError();
}

// mixed-radix sizes: NL lanes per block with N*NL <= 2048 (two 32 KB buffers)
template <int NL, int KIND>
int launch_mr(const FftArgs &a, const Fields &F, const RadixPlan &rp, hipStream_t s)
{
    constexpr int NS = 256 / NL, NMAX = 2048 / NL;
    const int chunks = (a.nlanes + NL - 1) / NL;
    const int blocks = (chunks + kXcds - 1) / kXcds * kXcds;
    const dim3 grid(blocks, F.nf), block(NL * NS);
    if (KIND == FWD64) hipLaunchKernelGGL((k_fft_forward_mr<NL, NS, NMAX>), grid, block, 0, s, a, F, rp);
    if (KIND == INV64) hipLaunchKernelGGL((k_fft_inverse_mr<NL, NS, NMAX>), grid, block, 0, s, a, F, rp);
    return (int)hipGetLastError();
}

template <int KIND>
int dispatch(int N, const FftArgs &a, const Fields &F, hipStream_t s)
{
    if (N & (N - 1)) {
        RadixPlan rp;
        if (N > 512 || !radix_plan(N, rp)) return (int)hipErrorInvalidValue;
        if (N <= 64) return launch_mr<32, KIND>(a, F, rp, s);
        if (N <= 128) return launch_mr<16, KIND>(a, F, rp, s);
        if (N <= 256) return launch_mr<8, KIND>(a, F, rp, s);
        return launch_mr<4, KIND>(a, F, rp, s);
    }
    switch (N) {
    case 1: return launch_one<1, 1, 256, KIND>(a, F, s);
    case 2: return launch_one<2, 1, 256, KIND>(a, F, s);
    case 4: return launch_one<4, 1, 256, KIND>(a, F, s);
    case 8: return launch_one<8, 1, 256, KIND>(a, F, s);
    case 16: return launch_one<16, 1, 256, KIND>(a, F, s);
    case 32: return launch_one<16, 2, 128, KIND>(a, F, s);
    case 64: return launch_one<16, 4, 64, KIND>(a, F, s);
    case 128: return launch_one<16, 8, 32, KIND>(a, F, s);
    case 256: return launch_one<16, 16, 16, KIND>(a, F, s);
    case 512: return launch_one<32, 16, 8, KIND>(a, F, s);
    default: return (int)hipErrorInvalidValue;
    }
}

FftArgs make_args(const chy_grid *g, int lanes)
{
    FftArgs a{};
    a.twL = g->d.twL;
    a.tw64 = g->d.tw64;
    a.twh = g->d.twh;
    a.dv = g->d.dv;
    a.J = g->J; a.K = g->K; a.L = g->L;
    a.JS = g->J + 2; a.KS = g->K + 2;
    a.lanes = lanes;
    a.nlanes = lanes == LANES_ALL ? g->J * g->K
             : lanes == LANES_INTERIOR ? (g->J - 1) * (g->K - 1) : g->J + g->K - 1;
    a.sf1 = 0.5 / (double)g->N;
    a.pi = g->pi;
    a.ys = g->pig4;
    return a;
}

}  // namespace

int launch_fft_forward_f(const chy_grid *g, const Fields &F, int coef_mode, int lanes, bool moments,
                         bool spectrum, hipStream_t s)
{
    FftArgs a = make_args(g, lanes);
    // non-null markers: bind() takes every field's own pointer
    a.rho = F.rho[0];
    a.phi = F.phi[0];
    a.coef_mode = coef_mode;
    a.mom = moments ? F.mom[0] : nullptr;
    a.spec = spectrum ? F.spec[0] : nullptr;
    return dispatch<FWD64>(g->N, a, F, s);
}

int launch_fft_forward(const chy_grid *g, const double *rho, const double *phi, double *coef,
                       int coef_mode, int lanes, bool moments, bool spectrum, hipStream_t s)
{
    const Fields F = one_field(g, rho, const_cast<double *>(phi), coef, 0);
    return launch_fft_forward_f(g, F, coef_mode, lanes, moments, spectrum, s);
}

int launch_fft_inverse_f(const chy_grid *g, const Fields &F, hipStream_t s)
{
    FftArgs a = make_args(g, LANES_ALL);
    a.spec = F.spec[0];
    a.phi_out = F.phi[0];
    a.ghosts = 1;
    if (zaxphi_from_samples(g)) a.zsamp = F.zsamp[0];  // non-null marker: bind() takes each field's
    return dispatch<INV64>(g->N, a, F, s);
}

bool fft_size_ok(int L)
{
    RadixPlan rp;
    return L >= 2 && L % 2 == 0 && L <= 1024 && radix_plan(L / 2, rp);
}

int launch_fft_inverse(const chy_grid *g, double *phi, hipStream_t s)
{
    const Fields F = one_field(g, nullptr, phi, nullptr, 0);
    return launch_fft_inverse_f(g, F, s);
}

}  // namespace chy
