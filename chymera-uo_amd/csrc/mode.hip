// Per-mode (r,z) solves of the potential solve.
//
// After the azimuthal FFT, every packed coefficient q (cosine part of order
// m=q for q<=N, sine part of order m=q-N for q>N) satisfies the 2-D system
// that POT3 hands to BLKTRI (pot/pot3.f:203-239, contract pot/blktri.f:3-56):
//
//   AN(k) X(j,k-1) + AM(j) X(j-1,k) + (BN(k)+BM_m(j)) X(j,k) + CM(j) X(j+1,k)
//   + CN(k) X(j,k+1) = Y(j,k),   j=1..M=JMAX-1 (r), k=1..NZ=KMAX-1 (z)
//
// with Y = 4 pi G PHI~ and the Dirichlet values moved to the right-hand side
// (pot/pot3.f:212-224).  The z operator tridiag(AN,BN,CN) does not depend on m;
// it is diagonalised once per grid (host, chy::build_grid), which turns each
// mode's system into NZ independent r-tridiagonal systems
//   (T_m + lambda_p) z_p = yhat_p,   yhat = Y W_f,   X = Z W_i^T.
// Three kernels:
//   k_zfwd    yhat[q][p][j] = sum_k W_f[k][p] Y_q[k][j]      (GEMM per mode)
//   k_rsolve  one wavefront per (m, p): both the cosine and sine right-hand
//             sides of order m, partitioned elimination in registers (each lane
//             owns R consecutive rows) + parallel cyclic reduction across the
//             64 lanes with cross-lane shuffles
//   k_zinv    X_q[k][j] = sum_p W_i[k][p] z[q][p][j], written back into the
//             spectrum (GEMM per mode)
#include "chypot_impl.h"

namespace chy {
namespace {

constexpr int TILE = 64;   // output tile edge
constexpr int KSTEP = 16;  // reduction step
constexpr int GEMM_THREADS = 256;

struct ModeArgs {
    double *spec;          // [L][K][J]
    double *yh;            // [L][PP][JP]
    const double *wf, *wi; // [NZ][PP]
    const double *lam;     // [PP]
    const double *am, *cm; // [M]
    const double *bm;      // [N+1][M]
    int J, K, L, N, M, NZ, PP, JP;
    double pig4, cnmax, cmmax;
};

// yhat[q][p][j] = sum_{k<NZ} wf[k][p] * Y_q[k][j], j < M, p < NZ (other p rows
// of the tile are written as 0).  Y_q is formed while loading (pot3.f:212-224):
//   Y = PIG4*S[k][j];  k=NZ-1: Y -= CNMAX*S[NZ][j];  j=M-1: Y -= CMMAX*S[k][M]
__global__ __launch_bounds__(GEMM_THREADS) void k_zfwd(ModeArgs a)
{
    __shared__ double As[KSTEP][TILE];      // [k][p]
    __shared__ double Bs[KSTEP][TILE + 1];  // [k][j]
    const int q = blockIdx.z;
    const int p0 = blockIdx.y * TILE, j0 = blockIdx.x * TILE;
    const int tid = threadIdx.x;
    const int tp = tid / 16, tj = tid % 16;
    const size_t JK = (size_t)a.J * a.K;
    const double *S = a.spec + (size_t)q * JK;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < a.NZ; k0 += KSTEP) {
        for (int e = tid; e < KSTEP * TILE; e += GEMM_THREADS) {
            const int kk = e / TILE, c = e % TILE;
            const int k = k0 + kk;
            const int p = p0 + c, j = j0 + c;
            As[kk][c] = (k < a.NZ && p < a.PP) ? a.wf[(size_t)k * a.PP + p] : 0.0;
            double y = 0.0;
            if (k < a.NZ && j < a.M) {
                y = a.pig4 * S[(size_t)k * a.J + j];
                if (k == a.NZ - 1) y -= a.cnmax * S[(size_t)a.NZ * a.J + j];
                if (j == a.M - 1) y -= a.cmmax * S[(size_t)k * a.J + a.M];
            }
            Bs[kk][c] = y;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KSTEP; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][tp * 4 + i];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][tj + 16 * c];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[i][c] = fma(av[i], bv[c], acc[i][c]);
        }
        __syncthreads();
    }
    double *Y = a.yh + (size_t)q * a.PP * a.JP;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int p = p0 + tp * 4 + i;
        if (p >= a.PP) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = j0 + tj + 16 * c;
            if (j < a.JP) Y[(size_t)p * a.JP + j] = (p < a.NZ && j < a.M) ? acc[i][c] : 0.0;
        }
    }
}

// spec[q][k][j] = sum_{p<NZ} wi[k][p] * z[q][p][j] for k < NZ, j < M.
__global__ __launch_bounds__(GEMM_THREADS) void k_zinv(ModeArgs a)
{
    __shared__ double As[KSTEP][TILE + 1];  // [p][k]
    __shared__ double Bs[KSTEP][TILE];      // [p][j]
    const int q = blockIdx.z;
    const int k0t = blockIdx.y * TILE, j0 = blockIdx.x * TILE;
    const int tid = threadIdx.x;
    const int tk = tid / 16, tj = tid % 16;
    const double *Z = a.yh + (size_t)q * a.PP * a.JP;
    double acc[4][4] = {};
    for (int pp0 = 0; pp0 < a.NZ; pp0 += KSTEP) {
        for (int e = tid; e < KSTEP * TILE; e += GEMM_THREADS) {
            // A tile: 64 k rows x 16 p, read along p
            const int r = e / KSTEP, c = e % KSTEP;
            const int k = k0t + r, p = pp0 + c;
            As[c][r] = (k < a.NZ && p < a.NZ) ? a.wi[(size_t)k * a.PP + p] : 0.0;
            // B tile: 16 p rows x 64 j, read along j
            const int pb = pp0 + e / TILE, jb = j0 + e % TILE;
            Bs[e / TILE][e % TILE] = (pb < a.NZ && jb < a.M) ? Z[(size_t)pb * a.JP + jb] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KSTEP; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][tk * 4 + i];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[kk][tj + 16 * c];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[i][c] = fma(av[i], bv[c], acc[i][c]);
        }
        __syncthreads();
    }
    double *S = a.spec + (size_t)q * a.J * a.K;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = k0t + tk * 4 + i;
        if (k >= a.NZ) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = j0 + tj + 16 * c;
            if (j < a.M) S[(size_t)k * a.J + j] = acc[i][c];
        }
    }
}

// One row of the reduced (per-lane) tridiagonal system.
struct Red {
    double al, be, ga;  // sub, diagonal, super coefficients
    double d0, d1;      // right-hand sides (cosine, sine)
};

__device__ __forceinline__ Red shfl_up_red(const Red &r, int s, int lane)
{
    Red o;
    o.al = __shfl_up(r.al, s, 64);
    o.be = __shfl_up(r.be, s, 64);
    o.ga = __shfl_up(r.ga, s, 64);
    o.d0 = __shfl_up(r.d0, s, 64);
    o.d1 = __shfl_up(r.d1, s, 64);
    if (lane < s) o = Red{0.0, 1.0, 0.0, 0.0, 0.0};
    return o;
}

__device__ __forceinline__ Red shfl_down_red(const Red &r, int s, int lane)
{
    Red o;
    o.al = __shfl_down(r.al, s, 64);
    o.be = __shfl_down(r.be, s, 64);
    o.ga = __shfl_down(r.ga, s, 64);
    o.d0 = __shfl_down(r.d0, s, 64);
    o.d1 = __shfl_down(r.d1, s, 64);
    if (lane + s >= 64) o = Red{0.0, 1.0, 0.0, 0.0, 0.0};
    return o;
}

// One wavefront solves (T_m + lambda_p) x = y for the cosine (q=m) and sine
// (q=N+m, 1<=m<=N-1) right-hand sides.  Lane i owns rows [iR, iR+R); rows >= M
// are identity rows.
template <int R>
__global__ __launch_bounds__(256) void k_rsolve(ModeArgs a)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sys = blockIdx.x * 4 + wave;
    if (sys >= (a.N + 1) * a.NZ) return;  // whole wavefront exits together
    const int m = sys / a.NZ, p = sys % a.NZ;
    const bool two = (m >= 1 && m <= a.N - 1);
    double *y0 = a.yh + ((size_t)m * a.PP + p) * a.JP;
    double *y1 = a.yh + ((size_t)(a.N + m) * a.PP + p) * a.JP;
    const double *bm = a.bm + (size_t)m * a.M;
    const double lam = a.lam[p];
    const int f = lane * R;

    double ca[R], cb[R], cc[R], d0[R], d1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = f + r;
        if (j < a.M) {
            ca[r] = a.am[j];
            cb[r] = bm[j] + lam;
            cc[r] = a.cm[j];
            d0[r] = y0[j];
            d1[r] = two ? y1[j] : 0.0;
        } else {
            ca[r] = 0.0; cb[r] = 1.0; cc[r] = 0.0; d0[r] = 0.0; d1[r] = 0.0;
        }
    }

    // local elimination of rows f..f+R-2 with the neighbouring interface
    // unknowns XL = x_{f-1} (last row of the previous lane) and XR = x_{f+R-1}
    // kept symbolic: x_i = D_i + P_i XL + Q_i XR
    double D0[R], D1[R], P[R], Q[R];
    Red red;
    if constexpr (R == 1) {
        red = Red{ca[0], cb[0], cc[0], d0[0], d1[0]};
    } else {
        double g[R], w0[R], w1[R], wp[R], wq[R];
#pragma unroll
        for (int i = 0; i < R - 1; ++i) {
            const double den = i == 0 ? cb[0] : cb[i] - ca[i] * g[i - 1];
            const double inv = 1.0 / den;
            g[i] = cc[i] * inv;
            const double rp = i == 0 ? -ca[0] : 0.0;
            const double rq = i == R - 2 ? -cc[R - 2] : 0.0;
            w0[i] = (d0[i] - (i ? ca[i] * w0[i - 1] : 0.0)) * inv;
            w1[i] = (d1[i] - (i ? ca[i] * w1[i - 1] : 0.0)) * inv;
            wp[i] = (rp - (i ? ca[i] * wp[i - 1] : 0.0)) * inv;
            wq[i] = (rq - (i ? ca[i] * wq[i - 1] : 0.0)) * inv;
        }
        D0[R - 2] = w0[R - 2]; D1[R - 2] = w1[R - 2]; P[R - 2] = wp[R - 2]; Q[R - 2] = wq[R - 2];
#pragma unroll
        for (int i = R - 3; i >= 0; --i) {
            D0[i] = w0[i] - g[i] * D0[i + 1];
            D1[i] = w1[i] - g[i] * D1[i + 1];
            P[i] = wp[i] - g[i] * P[i + 1];
            Q[i] = wq[i] - g[i] * Q[i + 1];
        }
        // first-row data of the next lane
        const double nD0 = __shfl_down(D0[0], 1, 64), nD1 = __shfl_down(D1[0], 1, 64);
        const double nP = __shfl_down(P[0], 1, 64), nQ = __shfl_down(Q[0], 1, 64);
        const double al = ca[R - 1], be = cb[R - 1], c = lane < 63 ? cc[R - 1] : 0.0;
        const double nPv = lane < 63 ? nP : 0.0, nQv = lane < 63 ? nQ : 0.0;
        const double nD0v = lane < 63 ? nD0 : 0.0, nD1v = lane < 63 ? nD1 : 0.0;
        red.al = al * P[R - 2];
        red.be = be + al * Q[R - 2] + c * nPv;
        red.ga = c * nQv;
        red.d0 = d0[R - 1] - al * D0[R - 2] - c * nD0v;
        red.d1 = d1[R - 1] - al * D1[R - 2] - c * nD1v;
    }

    // parallel cyclic reduction over the 64 lanes
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const Red lo = shfl_up_red(red, s, lane);
        const Red hi = shfl_down_red(red, s, lane);
        const double k1 = red.al / lo.be;
        const double k2 = red.ga / hi.be;
        Red nr;
        nr.al = -lo.al * k1;
        nr.ga = -hi.ga * k2;
        nr.be = red.be - lo.ga * k1 - hi.al * k2;
        nr.d0 = red.d0 - lo.d0 * k1 - hi.d0 * k2;
        nr.d1 = red.d1 - lo.d1 * k1 - hi.d1 * k2;
        red = nr;
    }
    const double xr0 = red.d0 / red.be, xr1 = red.d1 / red.be;

    // back-substitution into the lane's own rows
    double x0[R], x1[R];
    x0[R - 1] = xr0;
    x1[R - 1] = xr1;
    if constexpr (R > 1) {
        double xl0 = __shfl_up(xr0, 1, 64), xl1 = __shfl_up(xr1, 1, 64);
        if (lane == 0) { xl0 = 0.0; xl1 = 0.0; }
#pragma unroll
        for (int i = 0; i < R - 1; ++i) {
            x0[i] = D0[i] + P[i] * xl0 + Q[i] * xr0;
            x1[i] = D1[i] + P[i] * xl1 + Q[i] * xr1;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int j = f + r;
        if (j < a.JP) {
            y0[j] = j < a.M ? x0[r] : 0.0;
            if (two) y1[j] = j < a.M ? x1[r] : 0.0;
        }
    }
}

ModeArgs make_args(const chy_grid *g)
{
    ModeArgs a{};
    a.spec = g->d.spec;
    a.yh = g->d.yh;
    a.wf = g->d.wf;
    a.wi = g->d.wi;
    a.lam = g->d.lam;
    a.am = g->d.am;
    a.cm = g->d.cm;
    a.bm = g->d.bm;
    a.J = g->J; a.K = g->K; a.L = g->L; a.N = g->N;
    a.M = g->M; a.NZ = g->NZ; a.PP = g->PP; a.JP = g->JP;
    a.pig4 = g->pig4; a.cnmax = g->cnmax; a.cmmax = g->cmmax;
    return a;
}

}  // namespace

int launch_mode_solve(const chy_grid *g, hipStream_t s, Profiler *prof)
{
    ModeArgs a = make_args(g);
    dim3 gf((g->JP + TILE - 1) / TILE, (g->PP + TILE - 1) / TILE, g->L);
    if (prof) prof->begin(CHY_STAGE_ZFWD, s);
    hipLaunchKernelGGL(k_zfwd, gf, dim3(GEMM_THREADS), 0, s, a);
    if (prof) prof->end(s);
    const int nsys = (g->N + 1) * g->NZ;
    const int blocks = (nsys + 3) / 4;
    if (prof) prof->begin(CHY_STAGE_RSOLVE, s);
    switch (g->JP / 64) {
    case 1: hipLaunchKernelGGL(k_rsolve<1>, dim3(blocks), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_rsolve<2>, dim3(blocks), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_rsolve<4>, dim3(blocks), dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_rsolve<8>, dim3(blocks), dim3(256), 0, s, a); break;
    case 16: hipLaunchKernelGGL(k_rsolve<16>, dim3(blocks), dim3(256), 0, s, a); break;
    default: return (int)hipErrorInvalidValue;
    }
    if (prof) prof->end(s);
    dim3 gi((g->M + TILE - 1) / TILE, (g->NZ + TILE - 1) / TILE, g->L);
    if (prof) prof->begin(CHY_STAGE_ZINV, s);
    hipLaunchKernelGGL(k_zinv, gi, dim3(GEMM_THREADS), 0, s, a);
    if (prof) prof->end(s);
    return (int)hipGetLastError();
}

}  // namespace chy
