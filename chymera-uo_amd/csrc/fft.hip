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
//     x_n = a_0/2 + sum_{m=1}^{N-1} (a_m cos + b_m sin)(2 pi m n/L) + a_N/2 (-1)^n.
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

struct FftArgs {
    const double *rho;    // RHO (forward input on the interior)
    const double *phi;    // PHI (forward input on the boundary lanes)
    double *phi_out;      // PHI (inverse output)
    double *spec;         // [L][K][J]
    double *coef;         // COEF(JMAX+2,KMAX+2,LMAX/2,2) or null
    const double2 *twL;   // exp(2 pi i e/L)
    const double2 *tw64;  // exp(2 pi i e/64)
    int J, K, L, JS, KS;
    int nlanes;
    int coef_mode;
    double sf1, pi;
};

template <int N1, int N2, int NL>
__global__ __launch_bounds__(NL *N2) void k_fft_forward(FftArgs a)
{
    constexpr int N = N1 * N2;
    __shared__ double2 lds[N * NL];
    const int tid = threadIdx.x;
    const int lane = tid % NL;
    const int sub = tid / NL;
    const long lg = (long)blockIdx.x * NL + lane;
    const bool active = lg < a.nlanes;
    const int j = active ? (int)(lg % a.J) : 0;
    const int k = active ? (int)(lg / a.J) : 0;
    // interior lanes J<=JMAX, K<=KMAX carry the density (pot3.f:71-81)
    const double *src = (j <= a.J - 2 && k <= a.K - 2) ? a.rho : a.phi;
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);

    // phase 1: N1-point DFTs over z_{N2*n1+n2}, then twiddle W_N^{n2 k1}
    double2 v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        const int n = N2 * n1 + sub;
        double x0 = 0.0, x1 = 0.0;
        if (active) {
            x0 = src[base + plane * (size_t)(2 * n)];
            x1 = src[base + plane * (size_t)(2 * n + 1)];
        }
        v[n1] = make_double2(x0, x1);
    }
    dft_reg<N1, +1>(v, a.tw64);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        if (k1 > 0 && sub > 0) v[k1] = cmul(v[k1], a.twL[(2 * sub * k1) % a.L]);
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

    // phase 3: real split X_m = E_m + e^{2 pi i m/L} O_m, A_m = 2 Re X_m, B_m = 2 Im X_m
    auto xm = [&](int m) -> double2 {
        double2 zm = lds[(m % N) * NL + lane];
        double2 zc = lds[((N - m) % N) * NL + lane];
        zc.y = -zc.y;
        double2 e = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
        double2 d = csub(zm, zc);
        double2 o = make_double2(0.5 * d.y, -0.5 * d.x);
        return cadd(e, cmul(a.twL[m], o));
    };
    const size_t JK = (size_t)a.J * a.K;
    const size_t sidx = (size_t)k * a.J + j;
    const bool want_coef = a.coef && (a.coef_mode == CHY_COEF_FULL || (a.coef_mode == CHY_COEF_K2 && k == 0));
    double A0 = 0.0;
    if (want_coef) A0 = 2.0 * xm(0).x;
    for (int m = sub; m <= N; m += N2) {
        const double2 X = xm(m);
        const double A = 2.0 * X.x, B = 2.0 * X.y;
        a.spec[(size_t)m * JK + sidx] = A * a.sf1;
        if (m >= 1 && m <= N - 1) a.spec[(size_t)(N + m) * JK + sidx] = B * a.sf1;
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

template <int N1, int N2, int NL>
__global__ __launch_bounds__(NL *N2) void k_fft_inverse(FftArgs a)
{
    constexpr int N = N1 * N2;
    __shared__ double2 lds[N * NL];
    const int tid = threadIdx.x;
    const int lane = tid % NL;
    const int sub = tid / NL;
    const long lg = (long)blockIdx.x * NL + lane;
    const bool active = lg < a.nlanes;
    const int j = active ? (int)(lg % a.J) : 0;
    const int k = active ? (int)(lg / a.J) : 0;
    const size_t JK = (size_t)a.J * a.K;
    const size_t sidx = (size_t)k * a.J + j;

    // phase 0: C_0 = a_0/2, C_N = a_N/2 (both real, packed in slot 0),
    //          C_m = (a_m + i b_m)/2, m = 1..N-1
    for (int m = sub; m <= N; m += N2) {
        double av = 0.0, bv = 0.0;
        if (active) {
            av = a.spec[(size_t)m * JK + sidx];
            if (m >= 1 && m <= N - 1) bv = a.spec[(size_t)(N + m) * JK + sidx];
        }
        if (m == 0) lds[lane].x = 0.5 * av;
        else if (m == N) lds[lane].y = 0.5 * av;
        else lds[m * NL + lane] = make_double2(0.5 * av, 0.5 * bv);
    }
    __syncthreads();
    // phase 0b: Z_k = (C_k + conj C_{N-k}) + i e^{-2 pi i k/L} (C_k - conj C_{N-k})
    for (int kk = sub; kk <= N / 2; kk += N2) {
        if (kk == 0) {
            const double2 c = lds[lane];
            lds[lane] = make_double2(c.x + c.y, c.x - c.y);
        } else {
            const int kn = N - kk;
            const double2 ck = lds[kk * NL + lane], cn = lds[kn * NL + lane];
            const double2 U = make_double2(ck.x + cn.x, ck.y - cn.y);
            const double2 V = make_double2(ck.x - cn.x, ck.y + cn.y);
            double2 w = a.twL[kk];
            w.y = -w.y;
            const double2 wv = cmul(w, V);
            lds[kk * NL + lane] = make_double2(U.x - wv.y, U.y + wv.x);
            if (kn != kk) lds[kn * NL + lane] = make_double2(U.x + wv.y, -U.y + wv.x);
        }
    }
    __syncthreads();

    // phase 1: z_n = sum_k Z_k e^{-2 pi i k n/N}, split as the forward pass
    double2 v[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[n1] = lds[(N2 * n1 + sub) * NL + lane];
    dft_reg<N1, -1>(v, a.tw64);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        if (sub > 0) {
            double2 w = a.twL[(2 * sub * k1) % a.L];
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
    if (!active) return;
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int k1 = sub + N2 * r;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const int n = k1 + N1 * k2;
            a.phi_out[base + plane * (size_t)(2 * n)] = u[r][k2].x;
            a.phi_out[base + plane * (size_t)(2 * n + 1)] = u[r][k2].y;
        }
    }
}

template <int N1, int N2, int NL, bool FWD>
int launch_one(const FftArgs &a, hipStream_t s)
{
    const int blocks = (a.nlanes + NL - 1) / NL;
    if (FWD) hipLaunchKernelGGL((k_fft_forward<N1, N2, NL>), dim3(blocks), dim3(NL * N2), 0, s, a);
    else hipLaunchKernelGGL((k_fft_inverse<N1, N2, NL>), dim3(blocks), dim3(NL * N2), 0, s, a);
    return (int)hipGetLastError();
}

template <bool FWD>
int dispatch(int N, const FftArgs &a, hipStream_t s)
{
    switch (N) {
    case 1: return launch_one<1, 1, 256, FWD>(a, s);
    case 2: return launch_one<2, 1, 256, FWD>(a, s);
    case 4: return launch_one<4, 1, 256, FWD>(a, s);
    case 8: return launch_one<8, 1, 256, FWD>(a, s);
    case 16: return launch_one<16, 1, 256, FWD>(a, s);
    case 32: return launch_one<16, 2, 128, FWD>(a, s);
    case 64: return launch_one<16, 4, 64, FWD>(a, s);
    case 128: return launch_one<16, 8, 32, FWD>(a, s);
    case 256: return launch_one<16, 16, 16, FWD>(a, s);
    case 512: return launch_one<32, 16, 8, FWD>(a, s);
    default: return (int)hipErrorInvalidValue;
    }
}

FftArgs make_args(const chy_grid *g)
{
    FftArgs a{};
    a.spec = g->d.spec;
    a.twL = g->d.twL;
    a.tw64 = g->d.tw64;
    a.J = g->J; a.K = g->K; a.L = g->L;
    a.JS = g->J + 2; a.KS = g->K + 2;
    a.nlanes = g->J * g->K;
    a.sf1 = 0.5 / (double)g->N;
    a.pi = g->pi;
    return a;
}

}  // namespace

int launch_fft_forward(const chy_grid *g, const double *rho, const double *phi, double *coef,
                       int coef_mode, hipStream_t s)
{
    FftArgs a = make_args(g);
    a.rho = rho;
    a.phi = phi;
    a.coef = coef_mode == CHY_COEF_NONE ? nullptr : coef;
    a.coef_mode = coef_mode;
    return dispatch<true>(g->N, a, s);
}

int launch_fft_inverse(const chy_grid *g, double *phi, hipStream_t s)
{
    FftArgs a = make_args(g);
    a.phi_out = phi;
    return dispatch<false>(g->N, a, s);
}

}  // namespace chy
