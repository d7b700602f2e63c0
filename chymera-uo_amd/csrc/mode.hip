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
// Kernels:
//   k_rhs_edges  the boundary terms of Y on the K=KMAX row and J=JMAX column
//   z forward    yhat[q][p][j] = sum_k W_f[k][p] Y_q[k][j]      (GEMM per mode)
//   r solves     one wavefront per (m, p): both the cosine and sine right-hand
//                sides of order m, partitioned elimination in registers (each
//                lane owns R consecutive rows) + parallel cyclic reduction
//                across the 64 lanes with cross-lane shuffles
//   z inverse    X_q[k][j] = sum_p W_i[k][p] z[q][p][j], written back into the
//                spectrum (GEMM per mode)
#include <algorithm>

#include "chypot_impl.h"

namespace chy {
namespace {

struct ModeArgs {
    void *spec;            // [L][K][J] of TS (binary64, or binary32 for refinement)
    void *yh;              // [L][PP][JP] of TS
    const double *wf;      // [NZ][PP]  wf[k][p]
    const double *wi;      // [NZ][PP]  wi[p][k]
    const double *lam;     // [PP]
    const double *am, *cm; // [M]
    const double *bm;      // [N+1][M]
    int J, K, L, N, M, NZ, PP, JP;
    double pig4, cnmax, cmmax;
    const double *an_dev, *cn_dev;  // z coefficients AN, CN (spectral residual)
    int stat_step;                  // z inverse: refinement step whose statistics to collect (-1: none)
};

typedef double dx4 __attribute__((ext_vector_type(4)));

// z-transform GEMMs on the FP64 matrix cores, one 128x64 output tile per block:
//   FWD: yhat_q[p][j] = sum_k wf[k][p] Y_q[k][j]   rows p, cols j, reduction k
//   INV: X_q[k][j]    = sum_p wi[p][k] z_q[p][j]   rows k, cols j, reduction p
// Both weight tables are stored reduction-index major ([red][row]), so the A
// tile loads like the B tile: rows of contiguous doubles.  The four
// wavefronts own 32 rows each (2x4 tiles of 16x16); the next KT-deep slice of
// A and B is fetched into registers while the current one is multiplied.
// V_MFMA_F64_16X16X4F64 operands: A lane l holds A[l%16][l/16], B lane l holds
// B[l/16][l%16]; the result D[row][l%16] with row = MROW(l, v), v = 0..3.
constexpr int RT = 128, CT = 64, KT = 32;
constexpr int GEMM_THREADS = 256;
constexpr int A_PER = KT * RT / GEMM_THREADS;  // 16
constexpr int B_PER = KT * CT / GEMM_THREADS;  // 8

__device__ __forceinline__ int mrow(int lane, int v) { return (lane >> 4) + 4 * v; }

template <bool FWD, typename TS>
__global__ __launch_bounds__(GEMM_THREADS, 2) void k_zgemm(ModeArgs a, const Fields F)
{
    __shared__ double As[KT][RT + 1];
    __shared__ double Bs[KT][CT];
    const int q = blockIdx.z % a.L, f = blockIdx.z / a.L;  // mode, field
    a.spec = F.spec[f];
    a.yh = F.yh[f];
    const int r0 = blockIdx.y * RT, c0 = blockIdx.x * CT;
    const int tid = threadIdx.x;
    const int nred = a.NZ;
    const int nrow = FWD ? a.PP : a.NZ;
    const double *W = FWD ? a.wf : a.wi;
    TS *const specp = static_cast<TS *>(a.spec), *const yhp = static_cast<TS *>(a.yh);
    const TS *S = specp + (size_t)q * a.J * a.K;
    const TS *Z = yhp + (size_t)q * a.PP * a.JP;

    double ra[A_PER], rb[B_PER];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < A_PER; ++i) {
            const int e = tid + GEMM_THREADS * i;
            const int kk = e / RT, rr = e % RT;
            const int k = k0 + kk, r = r0 + rr;
            ra[i] = (k < nred && r < nrow) ? W[(size_t)k * a.PP + r] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < B_PER; ++i) {
            const int e = tid + GEMM_THREADS * i;
            const int kk = e / CT, cc = e % CT;
            const int k = k0 + kk, j = c0 + cc;
            double y = 0.0;
            if (k < nred && j < a.M) {
                y = FWD ? S[(size_t)k * a.J + j] : Z[(size_t)k * a.JP + j];
            }
            rb[i] = y;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < A_PER; ++i) {
            const int e = tid + GEMM_THREADS * i;
            As[e / RT][e % RT] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < B_PER; ++i) {
            const int e = tid + GEMM_THREADS * i;
            Bs[e / CT][e % CT] = rb[i];
        }
    };

    const int lane = tid & 63, wave = tid >> 6;
    const int rw = wave * 32, lr = lane & 15, lk = lane >> 4;
    dx4 acc[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[t][u] = dx4{0.0, 0.0, 0.0, 0.0};

    fetch(0);
    stash();
    __syncthreads();
    for (int k0 = 0; k0 < nred; k0 += KT) {
        const bool more = k0 + KT < nred;
        if (more) fetch(k0 + KT);
#pragma unroll
        for (int kk = 0; kk < KT; kk += 4) {
            const double a0 = As[kk + lk][rw + lr], a1 = As[kk + lk][rw + 16 + lr];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const double b = Bs[kk + lk][16 * u + lr];
                acc[0][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b, acc[0][u], 0, 0, 0);
                acc[1][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b, acc[1][u], 0, 0, 0);
            }
        }
        __syncthreads();
        if (more) {
            stash();
            __syncthreads();
        }
    }

#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = r0 + rw + 16 * t + mrow(lane, v);
                const int j = c0 + 16 * u + lr;
                if (FWD) {
                    if (r < a.PP && j < a.JP)
                        yhp[((size_t)q * a.PP + r) * a.JP + j] = (TS)((r < a.NZ && j < a.M) ? acc[t][u][v] : 0.0);
                } else if (r < a.NZ && j < a.M) {
                    specp[((size_t)q * a.K + r) * a.J + j] = (TS)acc[t][u][v];
                }
            }
}

// The same products for PP <= 128 (KMAX <= 128) with the weight table held in
// registers.  Persistent blocks (MINB per CU) walk output tiles of PPT rows x
// PCT columns (tile = (q, column block)); wave w owns rows 16w..16w+15, keeps
// its 16 x PPT slice of the weights as MFMA A operands for the whole run, and
// takes B operands from a double-buffered LDS tile that the block fills with
// the next tile's right-hand sides while the current one is multiplied.
typedef float fx4 __attribute__((ext_vector_type(4)));

// One 16x16x4 product on the matrix cores in the arithmetic of AT:
// V_MFMA_F64_16X16X4F64 or V_MFMA_F32_16X16X4F32 (same operand layout: A lane
// l holds A[l%16][l/16], B lane l holds B[l/16][l%16]).  The results differ in
// layout: element v of lane l is D[mrow_f64(l, v)][l%16] for binary64 and
// D[4 (l/16) + v][l%16] for binary32.
template <typename AT> struct Mfma;
template <> struct Mfma<double> {
    typedef dx4 acc;
    static __device__ __forceinline__ acc zero() { return dx4{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc mac(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
    typedef fx4 acc;
    static __device__ __forceinline__ acc zero() { return fx4{0.0f, 0.0f, 0.0f, 0.0f}; }
    static __device__ __forceinline__ acc mac(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Where the z inverse puts its result X (rows k < NZ, columns j < M):
//   EPI_SPEC  into the spectrum (type TO)
//   EPI_XSET  into the binary64 solution spectrum F.aux ([q][p][j] layout)
//   EPI_XADD  added to F.aux
//   EPI_XFIN  F.aux + X into the binary64 spectrum (the refinement's last step)
// AT: arithmetic of the products (binary64, or binary32 on the FP32 matrix
// cores for the refinement's correction solves); TI, TO: element types of the
// product's input and output (the forward transform reads the spectrum and
// writes eigen coordinates, the inverse the other way round).
template <bool FWD, int PPT, int PCT, int MINB, typename TI, typename TO, typename AT = double, int EPI = EPI_SPEC>
__global__ __launch_bounds__(PPT * 4, MINB * PPT / 64) void k_zgemm_reg(ModeArgs a, const Fields F, int col_tiles,
                                                                         int ntiles)
{
    typedef Mfma<AT> MM;
    // tile t: field t / tpf, then mode and column block within the field
    const int tpf = col_tiles * a.L;
    constexpr int NT = PPT * 4;            // threads (PPT/16 waves)
    constexpr int KS = PPT / 4;            // k-steps of the 16x16x4 product
    constexpr int U = PCT / 16;            // 16-column MFMA tiles per wave
    constexpr int BPER = PPT * PCT / NT;   // B elements fetched per thread
    __shared__ AT Bs[2][PPT][PCT];
    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const int nred = a.NZ;
    const int nrow = FWD ? a.PP : a.NZ;
    const double *W = FWD ? a.wf : a.wi;

    int tile = blockIdx.x;
    if (tile >= ntiles) return;  // the whole block leaves together

    // A operands: A[row][red] = W[red][row], row = 16w + lr, red = 4ks + lk
    AT ar[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int kr = 4 * ks + lk, r = 16 * w + lr;
        ar[ks] = (AT)((kr < nred && r < nrow) ? W[(size_t)kr * a.PP + r] : 0.0);
    }

    // fetch() only issues loads (clamped to valid addresses); stash(), after
    // the tile's products, masks and stores them, so no load result is
    // waited for inside the fetch.
    AT rb[BPER];
    const int cc = tid % PCT;  // this thread's column within every tile
    auto fetch = [&](int t) {
        const int f = t / tpf, tt = t % tpf;
        const int q = tt / col_tiles, j = (tt % col_tiles) * PCT + cc;
        const int jc = j < a.M ? j : a.M - 1;
        const TI *src = FWD ? static_cast<const TI *>(F.spec[f]) + (size_t)q * a.J * a.K
                            : static_cast<const TI *>(F.yh[f]) + (size_t)q * a.PP * a.JP;
        const int pitch = FWD ? a.J : a.JP;
#pragma unroll
        for (int i = 0; i < BPER; ++i) {
            const int kr = tid / PCT + (NT / PCT) * i;
            const int kc = kr < nred ? kr : nred - 1;
            rb[i] = (AT)src[(size_t)kc * pitch + jc];
        }
    };
    auto stash = [&](int buf, int t) {
        const int j = ((t % tpf) % col_tiles) * PCT + cc;
#pragma unroll
        for (int i = 0; i < BPER; ++i) {
            const int kr = tid / PCT + (NT / PCT) * i;
            Bs[buf][kr][cc] = (kr < nred && j < a.M) ? rb[i] : (AT)0;
        }
    };

    fetch(tile);
    stash(0, tile);
    __syncthreads();
    int buf = 0;
    double smax[2] = {0.0, 0.0};
    for (; tile < ntiles; tile += gridDim.x) {
        const int next = tile + gridDim.x;
        if (next < ntiles) fetch(next);
        // the solution spectrum this tile's results are added to, loaded
        // before the products so that the epilogue does not wait for it
        double xold[U][4];
        if constexpr (!FWD && (EPI == EPI_XADD || EPI == EPI_XFIN)) {
            const int f0 = tile / tpf, t0 = tile % tpf;
            const int q0 = t0 / col_tiles, c00 = (t0 % col_tiles) * PCT;
            const double *xf0 = static_cast<const double *>(F.aux[f0]);
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = 16 * w + MM::row(lane, v), j = c00 + 16 * u + lr;
                    const int rc = r < a.NZ ? r : a.NZ - 1, jc = j < a.M ? j : a.M - 1;
                    xold[u][v] = xf0[((size_t)q0 * a.PP + rc) * a.JP + jc];
                }
        }
        typename MM::acc acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = MM::zero();
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const AT b = Bs[buf][4 * ks + lk][16 * u + lr];
                acc[u] = MM::mac(ar[ks], b, acc[u]);
            }
        }
        const int f = tile / tpf, tt = tile % tpf;
        const int q = tt / col_tiles, c0 = (tt % col_tiles) * PCT;
        TO *const yhf = static_cast<TO *>(F.yh[f]);
        TO *const specf = static_cast<TO *>(F.spec[f]);
        double *const xf = static_cast<double *>(F.aux[f]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * w + MM::row(lane, v);
                const int j = c0 + 16 * u + lr;
                if (FWD) {
                    if (r < a.PP && j < a.JP)
                        yhf[((size_t)q * a.PP + r) * a.JP + j] = (TO)((r < a.NZ && j < a.M) ? acc[u][v] : (AT)0);
                } else if (r < a.NZ && j < a.M) {
                    const size_t xi = ((size_t)q * a.PP + r) * a.JP + j;
                    const double x = (double)acc[u][v];
                    double xn = x;
                    if constexpr (EPI == EPI_SPEC) specf[((size_t)q * a.K + r) * a.J + j] = (TO)acc[u][v];
                    if constexpr (EPI == EPI_XSET) xf[xi] = x;
                    if constexpr (EPI == EPI_XADD) {
                        xn = xold[u][v] + x;
                        xf[xi] = xn;
                    }
                    if constexpr (EPI == EPI_XFIN) {
                        xn = xold[u][v] + x;
                        static_cast<double *>(F.spec[f])[((size_t)q * a.K + r) * a.J + j] = xn;
                    }
                    if (!FWD && a.stat_step >= 0) {
                        smax[0] = fmax(smax[0], fabs(x));
                        smax[1] = fmax(smax[1], fabs(xn));
                    }
                }
            }
        // the refinement's step statistics: max |correction| and max |solution|
        // of the field, as the bit patterns of non-negative doubles
        if (!FWD && a.stat_step >= 0) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                smax[0] = fmax(smax[0], __shfl_xor(smax[0], o, 64));
                smax[1] = fmax(smax[1], __shfl_xor(smax[1], o, 64));
            }
            if (lane < 2) {
                unsigned long long *dst = F.rstat[f] + 2 * a.stat_step + lane;
                const unsigned long long bits = (unsigned long long)__double_as_longlong(smax[lane]);
                if (bits > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(dst, bits);
            }
            smax[0] = smax[1] = 0.0;
        }
        if (next < ntiles) stash(buf ^ 1, next);
        __syncthreads();
        buf ^= 1;
    }
}

// One row of the reduced (per-lane) tridiagonal system, in arithmetic AT.
template <typename AT>
struct Red {
    AT al, be, ga;  // sub, diagonal, super coefficients
    AT d0, d1;      // right-hand sides (cosine, sine)
};

template <typename AT>
__device__ __forceinline__ Red<AT> shfl_up_red(const Red<AT> &r, int s, int lane)
{
    Red<AT> o;
    o.al = __shfl_up(r.al, s, 64);
    o.be = __shfl_up(r.be, s, 64);
    o.ga = __shfl_up(r.ga, s, 64);
    o.d0 = __shfl_up(r.d0, s, 64);
    o.d1 = __shfl_up(r.d1, s, 64);
    if (lane < s) o = Red<AT>{0, 1, 0, 0, 0};
    return o;
}

template <typename AT>
__device__ __forceinline__ Red<AT> shfl_down_red(const Red<AT> &r, int s, int lane)
{
    Red<AT> o;
    o.al = __shfl_down(r.al, s, 64);
    o.be = __shfl_down(r.be, s, 64);
    o.ga = __shfl_down(r.ga, s, 64);
    o.d0 = __shfl_down(r.d0, s, 64);
    o.d1 = __shfl_down(r.d1, s, 64);
    if (lane + s >= 64) o = Red<AT>{0, 1, 0, 0, 0};
    return o;
}

// R consecutive elements at p (16-byte aligned) into registers of type AT,
// 16 bytes per load where the types allow
template <int R, typename TS, typename AT>
__device__ __forceinline__ void load_rows(const TS *p, AT (&d)[R])
{
    if constexpr (sizeof(TS) == 8 && R % 2 == 0) {
        const double2 *p2 = reinterpret_cast<const double2 *>(p);
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
            const double2 v = p2[i];
            d[2 * i] = (AT)v.x;
            d[2 * i + 1] = (AT)v.y;
        }
    } else if constexpr (sizeof(TS) == 4 && R % 4 == 0) {
        const float4 *p4 = reinterpret_cast<const float4 *>(p);
#pragma unroll
        for (int i = 0; i < R / 4; ++i) {
            const float4 v = p4[i];
            d[4 * i] = (AT)v.x; d[4 * i + 1] = (AT)v.y; d[4 * i + 2] = (AT)v.z; d[4 * i + 3] = (AT)v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) d[i] = (AT)p[i];
    }
}

template <int R, typename TS, typename AT>
__device__ __forceinline__ void store_rows(TS *p, const AT (&d)[R])
{
    if constexpr (sizeof(TS) == 8 && R % 2 == 0) {
        double2 *p2 = reinterpret_cast<double2 *>(p);
#pragma unroll
        for (int i = 0; i < R / 2; ++i) p2[i] = make_double2((double)d[2 * i], (double)d[2 * i + 1]);
    } else if constexpr (sizeof(TS) == 4 && R % 4 == 0) {
        float4 *p4 = reinterpret_cast<float4 *>(p);
#pragma unroll
        for (int i = 0; i < R / 4; ++i)
            p4[i] = make_float4((float)d[4 * i], (float)d[4 * i + 1], (float)d[4 * i + 2], (float)d[4 * i + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < R; ++i) p[i] = (TS)d[i];
    }
}

// One wavefront solves (T_m + lambda_p) x = y for two right-hand sides d0, d1
// (the cosine and sine parts of order m) in arithmetic AT.  Lane i owns rows
// [iR, iR+R); rows >= M are identity rows.  Inputs and outputs are the
// lane's R rows in registers.
template <int R, typename AT>
__device__ __forceinline__ void rsolve_rows(const ModeArgs &a, int m, int p, int lane, const AT (&d0)[R],
                                            const AT (&d1)[R], AT (&x0)[R], AT (&x1)[R])
{
    const AT lam = (AT)a.lam[p];
    const int f = lane * R;
    AT ca[R], cb[R], cc[R];
    load_rows<R>(a.am + f, ca);
    load_rows<R>(a.cm + f, cc);
    load_rows<R>(a.bm + (size_t)m * a.JP + f, cb);
#pragma unroll
    for (int r = 0; r < R; ++r) cb[r] = f + r < a.M ? cb[r] + lam : (AT)1;  // padding rows: x = 0

    // local elimination of rows f..f+R-2 with the neighbouring interface
    // unknowns XL = x_{f-1} (last row of the previous lane) and XR = x_{f+R-1}
    // kept symbolic: x_i = D_i + P_i XL + Q_i XR
    AT D0[R], D1[R], P[R], Q[R];
    Red<AT> red;
    if constexpr (R == 1) {
        red = Red<AT>{ca[0], cb[0], cc[0], d0[0], d1[0]};
    } else {
        AT g[R], w0[R], w1[R], wp[R], wq[R];
#pragma unroll
        for (int i = 0; i < R - 1; ++i) {
            const AT den = i == 0 ? cb[0] : cb[i] - ca[i] * g[i - 1];
            const AT inv = (AT)1 / den;
            g[i] = cc[i] * inv;
            const AT rp = i == 0 ? -ca[0] : (AT)0;
            const AT rq = i == R - 2 ? -cc[R - 2] : (AT)0;
            w0[i] = (d0[i] - (i ? ca[i] * w0[i - 1] : (AT)0)) * inv;
            w1[i] = (d1[i] - (i ? ca[i] * w1[i - 1] : (AT)0)) * inv;
            wp[i] = (rp - (i ? ca[i] * wp[i - 1] : (AT)0)) * inv;
            wq[i] = (rq - (i ? ca[i] * wq[i - 1] : (AT)0)) * inv;
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
        const AT nD0 = __shfl_down(D0[0], 1, 64), nD1 = __shfl_down(D1[0], 1, 64);
        const AT nP = __shfl_down(P[0], 1, 64), nQ = __shfl_down(Q[0], 1, 64);
        const AT al = ca[R - 1], be = cb[R - 1], c = lane < 63 ? cc[R - 1] : (AT)0;
        const AT nPv = lane < 63 ? nP : (AT)0, nQv = lane < 63 ? nQ : (AT)0;
        const AT nD0v = lane < 63 ? nD0 : (AT)0, nD1v = lane < 63 ? nD1 : (AT)0;
        red.al = al * P[R - 2];
        red.be = be + al * Q[R - 2] + c * nPv;
        red.ga = c * nQv;
        red.d0 = d0[R - 1] - al * D0[R - 2] - c * nD0v;
        red.d1 = d1[R - 1] - al * D1[R - 2] - c * nD1v;
    }

    // parallel cyclic reduction over the 64 lanes
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const Red<AT> lo = shfl_up_red(red, s, lane);
        const Red<AT> hi = shfl_down_red(red, s, lane);
        const AT k1 = red.al / lo.be;
        const AT k2 = red.ga / hi.be;
        Red<AT> nr;
        nr.al = -lo.al * k1;
        nr.ga = -hi.ga * k2;
        nr.be = red.be - lo.ga * k1 - hi.al * k2;
        nr.d0 = red.d0 - lo.d0 * k1 - hi.d0 * k2;
        nr.d1 = red.d1 - lo.d1 * k1 - hi.d1 * k2;
        red = nr;
    }
    const AT xr0 = red.d0 / red.be, xr1 = red.d1 / red.be;

    // back-substitution into the lane's own rows
    x0[R - 1] = xr0;
    x1[R - 1] = xr1;
    if constexpr (R > 1) {
        AT xl0 = __shfl_up(xr0, 1, 64), xl1 = __shfl_up(xr1, 1, 64);
        if (lane == 0) { xl0 = (AT)0; xl1 = (AT)0; }
#pragma unroll
        for (int i = 0; i < R - 1; ++i) {
            x0[i] = D0[i] + P[i] * xl0 + Q[i] * xr0;
            x1[i] = D1[i] + P[i] * xl1 + Q[i] * xr1;
        }
    }
}

// Standalone r solves: one wavefront per (m, p) over the z-eigen coordinates
// in yh, in place; the cosine (q=m) and sine (q=N+m, 1<=m<=N-1) right-hand
// sides together.  Orders m0 <= m < m1 (the refinement solves the lowest
// orders in binary64 and the rest in binary32); TS storage, AT arithmetic.
template <int R, typename TS, typename AT = double>
__global__ __launch_bounds__(256) void k_rsolve(ModeArgs a, const Fields F, int m0, int m1)
{
    a.yh = F.yh[blockIdx.y];  // field
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int sys = blockIdx.x * 4 + wave;
    if (sys >= (m1 - m0) * a.NZ) return;  // whole wavefront exits together
    const int m = m0 + sys / a.NZ, p = sys % a.NZ;
    const bool two = (m >= 1 && m <= a.N - 1);
    TS *y0 = static_cast<TS *>(a.yh) + ((size_t)m * a.PP + p) * a.JP;
    TS *y1 = static_cast<TS *>(a.yh) + ((size_t)(a.N + m) * a.PP + p) * a.JP;
    const int f = lane * R;
    // rows j >= M of yh are zero (written so by the z forward transform)
    AT d0[R], d1[R], x0[R], x1[R];
    load_rows<R>(y0 + f, d0);
    if (two) {
        load_rows<R>(y1 + f, d1);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) d1[r] = (AT)0;
    }
    rsolve_rows<R, AT>(a, m, p, lane, d0, d1, x0, x1);
    store_rows<R>(y0 + f, x0);
    if (two) store_rows<R>(y1 + f, x1);
}

// Y edge terms of pot3.f:212-224 on the spectrum (whose interior lanes the
// forward transform has already scaled by 4 pi G): the K=KMAX row loses
// CNMAX times the top boundary lane, the J=JMAX column CMMAX times the outer
// boundary lane.  One thread per (q, edge cell); the corner takes both.
__global__ __launch_bounds__(256) void k_rhs_edges(ModeArgs a, const Fields F)
{
    a.spec = F.spec[blockIdx.y];  // field
    const int ne = a.M + a.NZ - 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ne * a.L) return;
    const int q = (int)(t / ne), e = (int)(t % ne);
    double *S = static_cast<double *>(a.spec) + (size_t)q * a.J * a.K;
    if (e < a.M) {
        const size_t c = (size_t)(a.NZ - 1) * a.J + e;
        double y = S[c];
        y -= a.cnmax * S[(size_t)a.NZ * a.J + e];
        if (e == a.M - 1) y -= a.cmmax * S[c + 1];
        S[c] = y;
    } else {
        const size_t c = (size_t)(e - a.M) * a.J + (a.M - 1);
        S[c] -= a.cmmax * S[c + 1];
    }
}

ModeArgs make_args(const chy_grid *g)
{
    ModeArgs a{};
    a.wf = g->d.wf;
    a.wi = g->d.wi;
    a.lam = g->d.lam;
    a.am = g->d.am;
    a.cm = g->d.cm;
    a.bm = g->d.bm;
    a.J = g->J; a.K = g->K; a.L = g->L; a.N = g->N;
    a.M = g->M; a.NZ = g->NZ; a.PP = g->PP; a.JP = g->JP;
    a.pig4 = g->pig4; a.cnmax = g->cnmax; a.cmmax = g->cmmax;
    a.stat_step = -1;
    return a;
}

template <bool FWD, typename TI, typename TO, typename AT = double, int EPI = EPI_SPEC>
void launch_zgemm(const chy_grid *g, const ModeArgs &a, const Fields &F, hipStream_t s)
{
    // 64-column tiles, one persistent block per CU (128 KB / 64 KB of LDS)
    constexpr int PCT = 64;
    if (g->PP == 128 || g->PP == 64) {
        // binary32 forward products: two blocks per CU (half the LDS of a
        // binary64 block; the inverse's epilogue needs more than the 128
        // registers that would leave)
        constexpr int MB = (sizeof(AT) == 4 && FWD) ? 2 : 1;
        const int col_tiles = ((FWD ? g->JP : g->M) + PCT - 1) / PCT;
        const int ntiles = col_tiles * g->L * F.nf;
        const int blocks = std::min(ntiles, MB * g->ncu);
        if (g->PP == 128)
            hipLaunchKernelGGL((k_zgemm_reg<FWD, 128, PCT, MB, TI, TO, AT, EPI>), dim3(blocks), dim3(512), 0, s, a, F,
                               col_tiles, ntiles);
        else
            hipLaunchKernelGGL((k_zgemm_reg<FWD, 64, PCT, MB, TI, TO, AT, EPI>), dim3(blocks), dim3(256), 0, s, a, F,
                               col_tiles, ntiles);
    } else if (FWD) {
        dim3 gf((g->JP + CT - 1) / CT, (g->PP + RT - 1) / RT, g->L * F.nf);
        hipLaunchKernelGGL((k_zgemm<true, TO>), gf, dim3(GEMM_THREADS), 0, s, a, F);
    } else {
        dim3 gi((g->M + CT - 1) / CT, (g->NZ + RT - 1) / RT, g->L * F.nf);
        hipLaunchKernelGGL((k_zgemm<false, TO>), gi, dim3(GEMM_THREADS), 0, s, a, F);
    }
}

template <typename TS, typename AT = double>
int launch_rsolve(const chy_grid *g, const ModeArgs &a, const Fields &F, hipStream_t s, int m0, int m1)
{
    if (m1 <= m0) return (int)hipSuccess;
    const int nsys = (m1 - m0) * g->NZ;
    const dim3 grid((nsys + 3) / 4, F.nf), block(256);
    switch (g->JP / 64) {
    case 1: hipLaunchKernelGGL((k_rsolve<1, TS, AT>), grid, block, 0, s, a, F, m0, m1); break;
    case 2: hipLaunchKernelGGL((k_rsolve<2, TS, AT>), grid, block, 0, s, a, F, m0, m1); break;
    case 4: hipLaunchKernelGGL((k_rsolve<4, TS, AT>), grid, block, 0, s, a, F, m0, m1); break;
    case 8: hipLaunchKernelGGL((k_rsolve<8, TS, AT>), grid, block, 0, s, a, F, m0, m1); break;
    case 16: hipLaunchKernelGGL((k_rsolve<16, TS, AT>), grid, block, 0, s, a, F, m0, m1); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

}  // namespace

int launch_mode_solve_f(const chy_grid *g, const Fields &F, bool rhs_edges, hipStream_t s, Profiler *prof)
{
    ModeArgs a = make_args(g);
    if (g->mode_solver == CHY_MODES_CR) {
        if (prof) prof->begin(CHY_STAGE_CR, s);
        if (rhs_edges) {
            const long n = (long)(g->M + g->NZ - 1) * g->L;
            hipLaunchKernelGGL(k_rhs_edges, dim3((unsigned)((n + 255) / 256), F.nf), dim3(256), 0, s, a, F);
        }
        int rc = launch_cr_solve_f(g, F, s);
        if (prof) prof->end(s);
        return rc ? rc : (int)hipGetLastError();
    }
    if (prof) prof->begin(CHY_STAGE_ZFWD, s);
    if (rhs_edges) {
        const long n = (long)(g->M + g->NZ - 1) * g->L;
        hipLaunchKernelGGL(k_rhs_edges, dim3((unsigned)((n + 255) / 256), F.nf), dim3(256), 0, s, a, F);
    }
    launch_zgemm<true, double, double>(g, a, F, s);
    if (prof) prof->end(s);
    if (prof) prof->begin(CHY_STAGE_RSOLVE, s);
    int rc = launch_rsolve<double>(g, a, F, s, 0, g->N + 1);
    if (prof) prof->end(s);
    if (rc) return rc;
    if (prof) prof->begin(CHY_STAGE_ZINV, s);
    launch_zgemm<false, double, double>(g, a, F, s);
    if (prof) prof->end(s);
    return (int)hipGetLastError();
}

int launch_mode_solve(const chy_grid *g, bool rhs_edges, hipStream_t s, Profiler *prof)
{
    const Fields F = one_field(g, nullptr, nullptr, nullptr, 0);
    return launch_mode_solve_f(g, F, rhs_edges, s, prof);
}

int refine_supported(const chy_grid *g) { return g->PP == 128 || g->PP == 64; }

int launch_mode_solve32_f(const chy_grid *g, const Fields &Ff, const Fields &Fi, int m_fp64, bool rhs64, int epi,
                          int stat_step, hipStream_t s, Profiler *prof)
{
    if (!refine_supported(g)) return (int)hipErrorInvalidValue;
    ModeArgs a = make_args(g);
    const int m64 = std::max(-1, std::min(m_fp64, g->N));
    if (prof) prof->begin(CHY_STAGE_ZFWD, s);
    if (rhs64)
        launch_zgemm<true, double, float, float>(g, a, Ff, s);
    else
        launch_zgemm<true, float, float, float>(g, a, Ff, s);
    if (prof) prof->end(s);
    if (prof) prof->begin(CHY_STAGE_RSOLVE, s);
    int rc = launch_rsolve<float, double>(g, a, Ff, s, 0, m64 + 1);
    if (!rc) rc = launch_rsolve<float, float>(g, a, Ff, s, m64 + 1, g->N + 1);
    if (prof) prof->end(s);
    if (rc) return rc;
    a.stat_step = stat_step;
    if (prof) prof->begin(CHY_STAGE_ZINV, s);
    switch (epi) {
    case EPI_XSET: launch_zgemm<false, float, double, float, EPI_XSET>(g, a, Fi, s); break;
    case EPI_XADD: launch_zgemm<false, float, double, float, EPI_XADD>(g, a, Fi, s); break;
    case EPI_XFIN: launch_zgemm<false, float, double, float, EPI_XFIN>(g, a, Fi, s); break;
    default: launch_zgemm<false, float, double, float, EPI_SPEC>(g, a, Fi, s); break;
    }
    if (prof) prof->end(s);
    return (int)hipGetLastError();
}

__global__ __launch_bounds__(64) void k_zero_u64(const Fields F, int n)
{
    unsigned long long *p = F.rstat[blockIdx.x];
    for (int i = (int)threadIdx.x; i < n; i += 64) p[i] = 0ull;
}

int launch_zero_u64_f(const Fields &F, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(F.nf), dim3(64), 0, s, F, n);
    return (int)hipGetLastError();
}

int launch_rhs_edges_f(const chy_grid *g, const Fields &F, hipStream_t s)
{
    ModeArgs a = make_args(g);
    const long n = (long)(g->M + g->NZ - 1) * g->L;
    hipLaunchKernelGGL(k_rhs_edges, dim3((unsigned)((n + 255) / 256), F.nf), dim3(256), 0, s, a, F);
    return (int)hipGetLastError();
}

// Residual of the mode systems in the azimuthal spectrum, binary64:
//   R = B - (AN X(k-1) + AM X(j-1) + (BN + BM_m) X + CM X(j+1) + CN X(k+1))
// with B the right-hand side spectrum (F.spec, boundary terms applied) and X
// the binary64 solution spectrum (F.aux, [q][p][j]), rounded to binary32 into
// F.yh, a [q][k][j] array of pitch J (the next correction's right-hand side).
// The azimuthal DFT diagonalises the 7-point operator exactly (its L
// difference becomes 2 (cos m DTHETA - 1) C in BM_m, the axis coupling the
// ((-1)^m - 1) AMMIN term), so this is the residual of the physical-space
// operator, mode by mode.  One thread per (q, j); a sliding window down k.
__global__ __launch_bounds__(256) void k_spec_residual(ModeArgs a, const Fields F, const double *bn)
{
    const int q = blockIdx.y, fi = blockIdx.z;
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 256 + threadIdx.x;
    // every lane of a wavefront stays to the end (its neighbours read its
    // values); lanes past M compute on clamped columns and store nothing
    const bool live = j < a.M;
    const int jc = live ? j : a.M - 1;
    const int m = q <= a.N ? q : q - a.N;
    const double *B = static_cast<const double *>(F.spec[fi]) + (size_t)q * a.K * a.J + jc;
    const double *X = static_cast<const double *>(F.aux[fi]) + (size_t)q * a.PP * a.JP + jc;
    float *Rf = static_cast<float *>(F.yh[fi]) + (size_t)q * a.K * a.J + jc;
    const double amj = a.am[jc], cmj = a.cm[jc], bmj = a.bm[(size_t)m * a.JP + jc];  // AM(1) = CM(M) = 0
    // the wavefront's edge lanes take their outer neighbour from memory
    const bool jl = j > 0, jr = j + 1 < a.M;
    double xm = 0.0, x0 = X[0];
    for (int k = 0; k < a.NZ; ++k) {
        const double xp = k + 1 < a.NZ ? X[(size_t)(k + 1) * a.JP] : 0.0;
        double xl = __shfl_up(x0, 1, 64), xrt = __shfl_down(x0, 1, 64);
        if (lane == 0) xl = jl ? X[(size_t)k * a.JP - 1] : 0.0;
        if (lane == 63) xrt = jr ? X[(size_t)k * a.JP + 1] : 0.0;
        const double ax = a.an_dev[k] * xm + (bn[k] + bmj) * x0 + a.cn_dev[k] * xp + amj * xl + cmj * xrt;
        if (live) Rf[(size_t)k * a.J] = (float)(B[(size_t)k * a.J] - ax);
        xm = x0;
        x0 = xp;
    }
}

int launch_spec_residual_f(const chy_grid *g, const Fields &F, hipStream_t s)
{
    ModeArgs a = make_args(g);
    a.an_dev = g->d.rz_an;
    a.cn_dev = g->d.rz_cn;
    const dim3 grid((g->M + 255) / 256, g->L, F.nf);
    hipLaunchKernelGGL(k_spec_residual, grid, dim3(256), 0, s, a, F, g->d.rz_bn);
    return (int)hipGetLastError();
}

}  // namespace chy
