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
#include <type_traits>

#include "chypot_impl.h"
#include "lanes.h"

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

// 4 consecutive floats as one 16-byte store (p 16-byte aligned)
__device__ __forceinline__ void st4(float *p, const float (&x)[4])
{
    *reinterpret_cast<float4 *>(p) = make_float4(x[0], x[1], x[2], x[3]);
}

// Where the z inverse puts its result X (rows k < NZ, columns j < M):
//   EPI_SPEC  into the spectrum (type TO)
//   EPI_XSET  into the binary64 solution spectrum F.aux ([q][p][j] layout)
//   EPI_XADD  added to F.aux
//   EPI_XFIN  F.aux + X into the binary64 spectrum (the refinement's last step)
//   EPI_XSET32, EPI_XFIN32  the same with F.aux in binary32 (AT = float only:
//             X is then the binary32 sum of the products, stored exactly)
// AT: arithmetic of the products (binary64, or binary32 on the FP32 matrix
// cores for the refinement's correction solves); TI, TO: element types of the
// product's input and output (the forward transform reads the spectrum and
// writes eigen coordinates, the inverse the other way round).
template <bool FWD, int PPT, int PCT, int MINB, typename TI, typename TO, typename AT = double, int EPI = EPI_SPEC>
__global__ __launch_bounds__(PPT * 4, MINB * PPT / 64) void k_zgemm_reg(ModeArgs a, const Fields F, int col_tiles,
                                                                         int ntiles)
{
    typedef Mfma<AT> MM;
    constexpr bool X32 = EPI == EPI_XSET32 || EPI == EPI_XFIN32;
    static_assert(!X32 || sizeof(AT) == 4, "a binary32 solution spectrum holds binary32 sums only");
    typedef typename std::conditional<X32, float, double>::type XT;
    constexpr bool XSET = EPI == EPI_XSET || EPI == EPI_XSET32;
    constexpr bool XFIN = EPI == EPI_XFIN || EPI == EPI_XFIN32;
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
    // binary32 products with binary32 results run transposed (the LDS tile
    // as the first operand of V_MFMA_F32_16X16X4F32): a lane then holds 4
    // consecutive columns j of one row r = 16w + lr and stores them as one
    // 16-byte vector (a store instruction covers 16 rows x 64 bytes).  Forward
    // products 815 -> 756 and 741 -> 707 us, the X store 776 -> 764 us per
    // 32-field launch at 256x128x256.  Binary64 results stay in the other
    // layout, where 16 lanes write one row's 128 contiguous bytes: as two
    // 16-byte halves per lane they left holes in every store (999 -> 1228 us).
    // (The binary64 product's own layout has no run of columns in a lane.)
    constexpr bool TR = sizeof(AT) == 4 && ((FWD && sizeof(TO) == 4) || EPI == EPI_XSET32);

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
    // the tile's products, masks, converts and stores them, so no load result
    // is waited for inside the fetch (converting binary64 inputs to binary32
    // in the fetch made each load wait before the products: 750 -> 713 us
    // for the z forward of B per 32-field launch at 256x128x256).
    TI rb[BPER];
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
            rb[i] = src[(size_t)kc * pitch + jc];
        }
    };
    auto stash = [&](int buf, int t) {
        const int j = ((t % tpf) % col_tiles) * PCT + cc;
#pragma unroll
        for (int i = 0; i < BPER; ++i) {
            const int kr = tid / PCT + (NT / PCT) * i;
            Bs[buf][kr][cc] = (kr < nred && j < a.M) ? (AT)rb[i] : (AT)0;
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
        XT xold[U][4];
        if constexpr (!FWD && (EPI == EPI_XADD || XFIN)) {
            const int f0 = tile / tpf, t0 = tile % tpf;
            const int q0 = t0 / col_tiles, c00 = (t0 % col_tiles) * PCT;
            const XT *xf0 = static_cast<const XT *>(F.aux[f0]);
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
                if constexpr (TR)
                    acc[u] = MM::mac(b, ar[ks], acc[u]);
                else
                    acc[u] = MM::mac(ar[ks], b, acc[u]);
            }
        }
        const int f = tile / tpf, tt = tile % tpf;
        const int q = tt / col_tiles, c0 = (tt % col_tiles) * PCT;
        TO *const yhf = static_cast<TO *>(F.yh[f]);
        TO *const specf = static_cast<TO *>(F.spec[f]);
        XT *const xf = static_cast<XT *>(F.aux[f]);
        if constexpr (TR) {
            // rows r = 16w + lr, columns j0..j0+3; every vector lies inside
            // its row (JP is a multiple of the tile width), the stores of the
            // X rows stop at M
            const int r = 16 * w + lr;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j0 = c0 + 16 * u + 4 * lk;
                float o[4];
                if constexpr (FWD) {
                    if (r < a.PP) {
#pragma unroll
                        for (int v = 0; v < 4; ++v) o[v] = (r < a.NZ && j0 + v < a.M) ? acc[u][v] : 0.0f;
                        st4(reinterpret_cast<float *>(yhf) + ((size_t)q * a.PP + r) * a.JP + j0, o);
                    }
                } else if (r < a.NZ) {  // EPI_XSET32
                    const int n = a.M - j0;  // columns of the 4 inside the system
                    float *const p = reinterpret_cast<float *>(xf) + ((size_t)q * a.PP + r) * a.JP + j0;
#pragma unroll
                    for (int v = 0; v < 4; ++v) o[v] = acc[u][v];
                    if (n >= 4) {
                        st4(p, o);
                    } else {
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (v < n) p[v] = o[v];
                    }
                    if (a.stat_step >= 0) {
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (v < n) {
                                smax[0] = fmax(smax[0], fabs((double)o[v]));
                                smax[1] = smax[0];
                            }
                    }
                }
            }
        } else {
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
                    if constexpr (XSET) xf[xi] = (XT)x;
                    if constexpr (EPI == EPI_XADD) {
                        xn = (double)xold[u][v] + x;
                        xf[xi] = xn;
                    }
                    if constexpr (XFIN) {
                        xn = (double)xold[u][v] + x;
                        static_cast<double *>(F.spec[f])[((size_t)q * a.K + r) * a.J + j] = xn;
                    }
                    if (!FWD && a.stat_step >= 0) {
                        smax[0] = fmax(smax[0], fabs(x));
                        smax[1] = fmax(smax[1], fabs(xn));
                    }
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

// a / b in arithmetic AT: the binary64 division for binary64; for binary32
// (the refinement's correction solves, whose errors the next step removes)
// the product with the hardware reciprocal (1 ulp)
__device__ __forceinline__ double qdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// One row of the reduced (per-lane) tridiagonal system, in arithmetic AT.
template <typename AT>
struct Red {
    AT al, be, ga;  // sub, diagonal, super coefficients
    AT d0, d1;      // right-hand sides (cosine, sine)
};

// r, or the identity row where !keep, by selects: a branch here makes the
// wavefront wait for every shuffle in flight before it issues the next ones
template <typename AT>
__device__ __forceinline__ Red<AT> identity_unless(bool keep, const Red<AT> &r)
{
    return Red<AT>{keep ? r.al : (AT)0, keep ? r.be : (AT)1, keep ? r.ga : (AT)0, keep ? r.d0 : (AT)0,
                   keep ? r.d1 : (AT)0};
}

// the row s lanes below / above (distance 1 by DPP moves); identity rows
// past the wavefront's ends
template <typename AT>
__device__ __forceinline__ Red<AT> shfl_up_red(const Red<AT> &r, int s, int lane)
{
    Red<AT> o;
    if (s == 1) {
        o = Red<AT>{lane_prev(r.al), lane_prev(r.be), lane_prev(r.ga), lane_prev(r.d0), lane_prev(r.d1)};
    } else {
        o.al = __shfl_up(r.al, s, 64);
        o.be = __shfl_up(r.be, s, 64);
        o.ga = __shfl_up(r.ga, s, 64);
        o.d0 = __shfl_up(r.d0, s, 64);
        o.d1 = __shfl_up(r.d1, s, 64);
    }
    return identity_unless(lane >= s, o);
}

template <typename AT>
__device__ __forceinline__ Red<AT> shfl_down_red(const Red<AT> &r, int s, int lane)
{
    Red<AT> o;
    if (s == 1) {
        o = Red<AT>{lane_next(r.al), lane_next(r.be), lane_next(r.ga), lane_next(r.d0), lane_next(r.d1)};
    } else {
        o.al = __shfl_down(r.al, s, 64);
        o.be = __shfl_down(r.be, s, 64);
        o.ga = __shfl_down(r.ga, s, 64);
        o.d0 = __shfl_down(r.d0, s, 64);
        o.d1 = __shfl_down(r.d1, s, 64);
    }
    return identity_unless(lane + s < 64, o);
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
            const AT inv = qdiv((AT)1, den);
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
        const AT nD0 = lane_next(D0[0]), nD1 = lane_next(D1[0]);
        const AT nP = lane_next(P[0]), nQ = lane_next(Q[0]);
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
        const AT k1 = qdiv(red.al, lo.be);
        const AT k2 = qdiv(red.ga, hi.be);
        Red<AT> nr;
        nr.al = -lo.al * k1;
        nr.ga = -hi.ga * k2;
        nr.be = red.be - lo.ga * k1 - hi.al * k2;
        nr.d0 = red.d0 - lo.d0 * k1 - hi.d0 * k2;
        nr.d1 = red.d1 - lo.d1 * k1 - hi.d1 * k2;
        red = nr;
    }
    const AT xr0 = qdiv(red.d0, red.be), xr1 = qdiv(red.d1, red.be);

    // back-substitution into the lane's own rows
    x0[R - 1] = xr0;
    x1[R - 1] = xr1;
    if constexpr (R > 1) {
        AT xl0 = lane_prev(xr0), xl1 = lane_prev(xr1);
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

// r solves of grids whose systems do not fit a wavefront's registers
// (JMAX-1 > 1024): the Thomas algorithm, one thread per (m, p) with both
// right-hand sides, down and back up along r.  The systems are diagonally
// dominant (|BM + lambda_p| >= |AM| + |CM|), so no pivoting is needed.  The
// reciprocal pivots 1/(BM_m(j) + lambda_p - AM(j) g(j-1)), g(j) = CM(j) times
// that, depend on the grid alone and are tabulated once per grid in rinv
// ([m][p][j], binary64; k_thomas_pivots).
__global__ __launch_bounds__(256) void k_thomas_pivots(ModeArgs a, double *rinv)
{
    const int sys = blockIdx.x * 256 + threadIdx.x;
    if (sys >= (a.N + 1) * a.NZ) return;
    const int m = sys / a.NZ, p = sys % a.NZ;
    const double lam = a.lam[p];
    const double *bm = a.bm + (size_t)m * a.JP;
    double *out = rinv + (size_t)sys * a.JP;
    double g = 0.0;
    for (int j = 0; j < a.M; ++j) {
        const double inv = 1.0 / ((bm[j] + lam) - a.am[j] * g);
        out[j] = inv;
        g = a.cm[j] * inv;
    }
}

template <typename TS, typename AT>
__global__ __launch_bounds__(256) void k_rsolve_thomas(ModeArgs a, const Fields F, int m0, int m1,
                                                       const double *rinv)
{
    const int sys = blockIdx.x * 256 + threadIdx.x;
    if (sys >= (m1 - m0) * a.NZ) return;
    const int m = m0 + sys / a.NZ, p = sys % a.NZ;
    const bool two = (m >= 1 && m <= a.N - 1);
    TS *y0 = static_cast<TS *>(F.yh[blockIdx.y]) + ((size_t)m * a.PP + p) * a.JP;
    TS *y1 = static_cast<TS *>(F.yh[blockIdx.y]) + ((size_t)(a.N + m) * a.PP + p) * a.JP;
    const double *inv = rinv + ((size_t)m * a.NZ + p) * a.JP;
    AT d0 = (AT)0, d1 = (AT)0;
    for (int j = 0; j < a.M; ++j) {
        const AT iv = (AT)inv[j], aj = (AT)a.am[j];
        d0 = ((AT)y0[j] - aj * d0) * iv;
        y0[j] = (TS)d0;
        if (two) {
            d1 = ((AT)y1[j] - aj * d1) * iv;
            y1[j] = (TS)d1;
        }
    }
    AT x0 = (AT)0, x1 = (AT)0;
    for (int j = a.M - 1; j >= 0; --j) {
        const AT gj = (AT)(a.cm[j] * inv[j]);
        x0 = (AT)y0[j] - gj * x0;
        y0[j] = (TS)x0;
        if (two) {
            x1 = (AT)y1[j] - gj * x1;
            y1[j] = (TS)x1;
        }
    }
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
    // 64-column tiles, one persistent block per CU (128 KB / 64 KB of LDS);
    // the last binary32 inverse (binary32 X read, the binary64 spectrum
    // written) in 32-column tiles, whose epilogue fits two blocks per CU
    // (993 against 1147 us per 32-field launch at 256x128x256)
    constexpr int PCT = EPI == EPI_XFIN32 ? 32 : 64;
    if (g->PP == 128 || g->PP == 64) {
        // binary32 forward products and binary32-X inverses: two blocks per
        // CU (half the LDS of a binary64 block; the binary64-X inverses'
        // epilogues need more than the 128 registers that would leave)
        constexpr int MB = (sizeof(AT) == 4 && (FWD || EPI == EPI_XSET32 || EPI == EPI_XFIN32)) ? 2 : 1;
        const int col_tiles = ((FWD ? g->JP : g->M) + PCT - 1) / PCT;
        const int ntiles = col_tiles * g->L * F.nf;
        const int blocks = std::min(ntiles, MB * g->ncu);
        if (g->PP == 128)
            hipLaunchKernelGGL((k_zgemm_reg<FWD, 128, PCT, MB, TI, TO, AT, EPI>), dim3(blocks), dim3(512), 0, s, a, F,
                               col_tiles, ntiles);
        else
            hipLaunchKernelGGL((k_zgemm_reg<FWD, 64, PCT, MB, TI, TO, AT, EPI>), dim3(blocks), dim3(256), 0, s, a, F,
                               col_tiles, ntiles);
    } else if (g->PP == 256 && sizeof(AT) == 4) {
        // binary32 products of KMAX = 256 (the refined solve): 16 wavefronts
        // of 16 rows, 16-column tiles (wider tiles spill the 64 A operands'
        // neighbours out of the 128 registers a wavefront has), one block per CU
        if constexpr (sizeof(AT) == 4) {
            constexpr int PCT2 = 16;
            const int col_tiles = ((FWD ? g->JP : g->M) + PCT2 - 1) / PCT2;
            const int ntiles = col_tiles * g->L * F.nf;
            const int blocks = std::min(ntiles, g->ncu);
            hipLaunchKernelGGL((k_zgemm_reg<FWD, 256, PCT2, 1, TI, TO, AT, EPI>), dim3(blocks), dim3(1024), 0, s, a,
                               F, col_tiles, ntiles);
        }
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
    default:
        if (!g->d.rinv) return (int)hipErrorInvalidValue;
        hipLaunchKernelGGL((k_rsolve_thomas<TS, AT>), dim3((nsys + 255) / 256, F.nf), block, 0, s, a, F, m0, m1,
                           g->d.rinv);
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

int refine_supported(const chy_grid *g) { return g->PP == 256 || g->PP == 128 || g->PP == 64; }

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
    case EPI_XSET32: launch_zgemm<false, float, double, float, EPI_XSET32>(g, a, Fi, s); break;
    case EPI_XFIN32: launch_zgemm<false, float, double, float, EPI_XFIN32>(g, a, Fi, s); break;
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
// the solution spectrum (F.aux, [q][p][j], binary64, or binary32 when XT is:
// the first iterate of a 2-step solve), rounded to binary32 into
// F.yh, a [q][k][j] array of pitch J (the next correction's right-hand side).
// The azimuthal DFT diagonalises the 7-point operator exactly (its L
// difference becomes 2 (cos m DTHETA - 1) C in BM_m, the axis coupling the
// ((-1)^m - 1) AMMIN term), so this is the residual of the physical-space
// operator, mode by mode.  One thread per (q, j) and block of RROWS rows k:
// blockIdx.x is the row block, so the blocks in flight at once cover
// consecutive rows of few planes (one block walking every row of its plane
// kept a separate stream open in each of thousands of planes); every load of
// the block's rows is issued before the first is used.
constexpr int RROWS = 8;
template <typename XT>
__global__ __launch_bounds__(256) void k_spec_residual(ModeArgs a, const Fields F, const double *bn)
{
    const int q = blockIdx.y, fi = blockIdx.z;
    const int nkb = (a.NZ + RROWS - 1) / RROWS;  // row blocks per column block
    const int k0 = (int)(blockIdx.x % nkb) * RROWS;
    const int lane = threadIdx.x & 63;
    const int j = (int)(blockIdx.x / nkb) * 256 + threadIdx.x;
    // every lane of a wavefront stays to the end (its neighbours read its
    // values); lanes past M compute on clamped columns and store nothing
    const bool live = j < a.M;
    const int jc = live ? j : a.M - 1;
    const int m = q <= a.N ? q : q - a.N;
    const int NZ = a.NZ;
    const double *B = static_cast<const double *>(F.spec[fi]) + (size_t)q * a.K * a.J + jc;
    const XT *X = static_cast<const XT *>(F.aux[fi]) + (size_t)q * a.PP * a.JP + jc;
    float *Rf = static_cast<float *>(F.yh[fi]) + (size_t)q * a.K * a.J + jc;
    const double amj = a.am[jc], cmj = a.cm[jc], bmj = a.bm[(size_t)m * a.JP + jc];  // AM(1) = CM(M) = 0
    // the wavefront's edge lanes take their outer neighbour from memory (lane
    // 0 the left one, lane 63 the right one, the others their own column)
    const bool jl = j > 0, jr = j + 1 < a.M;
    const int eoff = (lane == 0 && jl) ? -1 : ((lane == 63 && jr) ? 1 : 0);
    // X rows k0-1 .. k0+RROWS, B rows k0 .. k0+RROWS-1, the edge neighbours:
    // every load at a clamped row, issued before any is used (a load under a
    // branch waits inside it when its value is converted there), then the
    // rows outside 0..NZ-1 selected away
    XT xv[RROWS + 2], ev[RROWS];
    double bs[RROWS];
#pragma unroll
    for (int i = 0; i < RROWS + 2; ++i) {
        const int k = k0 - 1 + i, kc = k < 0 ? 0 : (k < NZ ? k : NZ - 1);
        xv[i] = X[(size_t)kc * a.JP];
    }
#pragma unroll
    for (int i = 0; i < RROWS; ++i) {
        const int k = k0 + i, kc = k < NZ ? k : NZ - 1;
        bs[i] = B[(size_t)kc * a.J];
        ev[i] = X[(size_t)kc * a.JP + eoff];
    }
    double xs[RROWS + 2];
#pragma unroll
    for (int i = 0; i < RROWS + 2; ++i) {
        const int k = k0 - 1 + i;
        xs[i] = (k >= 0 && k < NZ) ? (double)xv[i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < RROWS; ++i) {
        const int k = k0 + i;
        if (k >= NZ) break;  // the same k for the whole block
        const double xm = xs[i], xc = xs[i + 1], xp = xs[i + 2];
        double xl = lane_prev(xc), xrt = lane_next(xc);
        if (lane == 0) xl = jl ? (double)ev[i] : 0.0;
        if (lane == 63) xrt = jr ? (double)ev[i] : 0.0;
        const double ax = a.an_dev[k] * xm + (bn[k] + bmj) * xc + a.cn_dev[k] * xp + amj * xl + cmj * xrt;
        if (live) Rf[(size_t)k * a.J] = (float)(bs[i] - ax);
    }
}

namespace {

// ---------------------------------------------------------------------------
// The refined solve's mode work for one mode plane in one kernel.
//
// One block per plane q of one field.
// The binary32 plane, PP rows x JP columns, lives in LDS for the whole plane
// and is rewritten in place by each phase; nothing but the right-hand side B
// (read twice) and the result X (written once) goes through HBM:
//   1. z forward of B (binary64 from the spectrum, in 32-row chunks that are
//      loaded while the previous chunk is multiplied):  Y = W_f^T B
//   2. r solves (T_m + lambda_p) y_p = Y_p: one lane per row p with the
//      grid's LU factors (binary64 solves of the orders m <= m64: one
//      wavefront per row, parallel cyclic reduction)
//   3. z inverse X1 = W_i Y (binary32)
// then, for two steps:
//   4. the binary64 residual R = B - A X1 of the plane's system, its stencil
//      read from X1 staged in LDS, rounded to binary32 (X1 also goes to a
//      binary32 scratch plane, F.aux, [q][PP][JP])
//   5. 1..3 again on R: the correction C
//   6. X = X1 + C in binary64 into the spectrum.
// With one step, X1 is the result.  HBM traffic per plane: B twice, X1 in
// binary32 out and back, X out.  The products run on the FP32 matrix
// cores (V_MFMA_F32_16X16X4F32): wave w owns output rows 16w..16w+15 and
// all JP columns, A operands from the binary32 weight tables, B operands
// from the LDS plane.  The r solves of the orders m <=
// m64 run in binary64 arithmetic, the others in binary32.  An LDS row is JP +
// 17 floats: the 64 lanes of a per-lane r solve read one column from 64
// banks, and a B operand's four rows meet at most two to a bank.  A block
// needs PP (JP + 17) 4 bytes of LDS: 140 KB at PP = 128, JP = 256.

// One r-tridiagonal system over the 64 lanes of a wavefront, lane i owning
// rows [iR, iR+R): cb already holds the diagonal (1 on padding rows, where
// d = 0); partitioned elimination in the lane, then parallel cyclic
// reduction of the lanes' interface rows.
template <int R, typename AT>
__device__ __forceinline__ void rsolve1(const AT (&ca)[R], const AT (&cb)[R], const AT (&cc)[R], int lane,
                                        const AT (&d)[R], AT (&x)[R])
{
    AT D[R], P[R], Q[R];
    AT al, be, ga, dd;
    if constexpr (R == 1) {
        al = ca[0]; be = cb[0]; ga = cc[0]; dd = d[0];
    } else {
        AT g[R], w[R], wp[R], wq[R];
#pragma unroll
        for (int i = 0; i < R - 1; ++i) {
            const AT den = i == 0 ? cb[0] : cb[i] - ca[i] * g[i - 1];
            const AT inv = qdiv((AT)1, den);
            g[i] = cc[i] * inv;
            const AT rp = i == 0 ? -ca[0] : (AT)0;
            const AT rq = i == R - 2 ? -cc[R - 2] : (AT)0;
            w[i] = (d[i] - (i ? ca[i] * w[i - 1] : (AT)0)) * inv;
            wp[i] = (rp - (i ? ca[i] * wp[i - 1] : (AT)0)) * inv;
            wq[i] = (rq - (i ? ca[i] * wq[i - 1] : (AT)0)) * inv;
        }
        D[R - 2] = w[R - 2]; P[R - 2] = wp[R - 2]; Q[R - 2] = wq[R - 2];
#pragma unroll
        for (int i = R - 3; i >= 0; --i) {
            D[i] = w[i] - g[i] * D[i + 1];
            P[i] = wp[i] - g[i] * P[i + 1];
            Q[i] = wq[i] - g[i] * Q[i + 1];
        }
        const AT nD = lane_next(D[0]), nP = lane_next(P[0]), nQ = lane_next(Q[0]);
        const bool last = lane == 63;
        const AT c = last ? (AT)0 : cc[R - 1];
        al = ca[R - 1] * P[R - 2];
        be = cb[R - 1] + ca[R - 1] * Q[R - 2] + c * (last ? (AT)0 : nP);
        ga = c * (last ? (AT)0 : nQ);
        dd = d[R - 1] - ca[R - 1] * D[R - 2] - c * (last ? (AT)0 : nD);
    }
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        AT lal, lbe, lga, ld, hal, hbe, hga, hd;
        if (s == 1) {
            lal = lane_prev(al); lbe = lane_prev(be); lga = lane_prev(ga); ld = lane_prev(dd);
            hal = lane_next(al); hbe = lane_next(be); hga = lane_next(ga); hd = lane_next(dd);
        } else {
            lal = __shfl_up(al, s, 64); lbe = __shfl_up(be, s, 64); lga = __shfl_up(ga, s, 64);
            ld = __shfl_up(dd, s, 64);
            hal = __shfl_down(al, s, 64); hbe = __shfl_down(be, s, 64); hga = __shfl_down(ga, s, 64);
            hd = __shfl_down(dd, s, 64);
        }
        // identity rows past the ends, by selects (see identity_unless)
        const bool kl = lane >= s, kh = lane + s < 64;
        lal = kl ? lal : (AT)0; lbe = kl ? lbe : (AT)1; lga = kl ? lga : (AT)0; ld = kl ? ld : (AT)0;
        hal = kh ? hal : (AT)0; hbe = kh ? hbe : (AT)1; hga = kh ? hga : (AT)0; hd = kh ? hd : (AT)0;
        const AT k1 = qdiv(al, lbe), k2 = qdiv(ga, hbe);
        const AT nal = -lal * k1, nga = -hga * k2;
        const AT nbe = be - lga * k1 - hal * k2;
        const AT nd = dd - ld * k1 - hd * k2;
        al = nal; ga = nga; be = nbe; dd = nd;
    }
    const AT xr = qdiv(dd, be);
    x[R - 1] = xr;
    if constexpr (R > 1) {
        AT xl = lane_prev(xr);
        if (lane == 0) xl = (AT)0;
#pragma unroll
        for (int i = 0; i < R - 1; ++i) x[i] = D[i] + P[i] * xl + Q[i] * xr;
    }
}

// The r solves of one plane in LDS (rows p < NZ), NS systems at a time per
// wavefront (independent chains, so each one's shuffle and reciprocal
// latency hides behind the others'); arithmetic AT.
template <int R, int LP, typename AT, int NS>
__device__ __forceinline__ void plane_rsolve(const ModeArgs &a, float *S, int m, int w, int lane)
{
    const int f = lane * R;
    AT ca[R], cc[R], cb0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        ca[r] = (AT)a.am[f + r];
        cc[r] = (AT)a.cm[f + r];
        cb0[r] = (AT)a.bm[(size_t)m * a.JP + f + r];
    }
#pragma unroll 1
    for (int i = 0; i < 16; i += NS) {
        const int p0 = 16 * w + i;
        if (p0 >= a.NZ) break;  // rows past NZ are zero and stay zero (wave-uniform)
        AT d[NS][R], x[NS][R], cb[NS][R];
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            const bool on = p0 + t < a.NZ;
            const AT lam = on ? (AT)a.lam[p0 + t] : (AT)0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                d[t][r] = on ? (AT)S[(p0 + t) * LP + f + r] : (AT)0;
                cb[t][r] = f + r < a.M ? cb0[r] + lam : (AT)1;
            }
        }
#pragma unroll
        for (int t = 0; t < NS; ++t) rsolve1<R, AT>(ca, cb[t], cc, lane, d[t], x[t]);
#pragma unroll
        for (int t = 0; t < NS; ++t)
            if (p0 + t < a.NZ)
#pragma unroll
                for (int r = 0; r < R; ++r) S[(p0 + t) * LP + f + r] = (float)x[t][r];
    }
}

// The binary32 r solves of one plane in LDS by one lane per system: lane l
// of wavefront w solves row p = 64 w + l (p < NZ) with the LU factors of
// T_m + lambda_p prepared once per grid, fac = (l_j, 1/pivot_j) per row j:
//   forward   D_j = d_j - l_j D_{j-1}
//   backward  x_j = (D_j - c_j x_{j+1}) / pivot_j
// over all JP rows (rows past M carry l = 0, 1/pivot = 1, c = 0, d = 0).
// Each step's chain is one FMA; the other products and every load are off
// the chain, and the loads of the next TB rows are issued before the current
// TB rows are computed.  The sweeps need no cross-lane traffic; the row
// pitch LP = JP + 17 puts the 64 lanes' elements of one column in 64
// different banks.  cS: c_j (CM) in LDS.
template <int LP, int JP>
__device__ __forceinline__ void plane_thomas(float *S, const float *cS, const float *fac, int PPg, int NZ, int w,
                                             int lane)
{
    const int p = 64 * w + lane;
    if (p >= NZ) return;
    float *row = S + p * LP;
    const float2 *fr = reinterpret_cast<const float2 *>(fac) + p;  // row j at fr[j PP]
    constexpr int TB = 16, NB = JP / TB;
    float2 fc[TB], fn[TB];
    float dc[TB], dn[TB];
    auto load = [&](float2 (&f)[TB], float (&d)[TB], int j0) {
#pragma unroll
        for (int i = 0; i < TB; ++i) {
            f[i] = fr[(size_t)(j0 + i) * PPg];
            d[i] = row[j0 + i];
        }
    };
    float D = 0.0f;
    load(fc, dc, 0);
#pragma unroll 1
    for (int b = 0; b < NB; ++b) {
        if (b + 1 < NB) load(fn, dn, (b + 1) * TB);
#pragma unroll
        for (int i = 0; i < TB; ++i) {
            D = __builtin_fmaf(-fc[i].x, D, dc[i]);
            row[b * TB + i] = D;
        }
#pragma unroll
        for (int i = 0; i < TB; ++i) { fc[i] = fn[i]; dc[i] = dn[i]; }
    }
    float x = 0.0f;
    load(fc, dc, JP - TB);
#pragma unroll 1
    for (int b = NB - 1; b >= 0; --b) {
        if (b > 0) load(fn, dn, (b - 1) * TB);
#pragma unroll
        for (int i = TB - 1; i >= 0; --i) {
            const int j = b * TB + i;
            const float inv = fc[i].y;
            x = __builtin_fmaf(-cS[j] * inv, x, dc[i] * inv);
            row[j] = x;
        }
#pragma unroll
        for (int i = 0; i < TB; ++i) { fc[i] = fn[i]; dc[i] = dn[i]; }
    }
}

// The LU factors of T_m + lambda_p for every order m <= N and z eigenvalue
// p < NZ, binary64 arithmetic rounded once: fac[((m JP + j) PP + p)] =
// (l_j, 1/pivot_j) with pivot_0 = b_0, l_j = a_j / pivot_{j-1}, pivot_j =
// b_j - l_j c_{j-1}, b_j = BM_m(j) + lambda_p; rows j >= M: (0, 1).  One
// thread per (m, p); p fastest, so the lanes of a per-lane solve read one
// row j of 64 systems together.
__global__ __launch_bounds__(256) void k_thomas_factors(ModeArgs a, float *fac)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (a.N + 1) * a.NZ) return;
    const int m = t / a.NZ, p = t % a.NZ;
    const double lam = a.lam[p];
    float2 *f = reinterpret_cast<float2 *>(fac) + (size_t)m * a.JP * a.PP + p;
    double piv = 1.0;
    for (int j = 0; j < a.JP; ++j) {
        if (j >= a.M) {
            f[(size_t)j * a.PP] = make_float2(0.0f, 1.0f);
            continue;
        }
        const double l = j > 0 ? a.am[j] / piv : 0.0;
        piv = a.bm[(size_t)m * a.JP + j] + lam - (j > 0 ? l * a.cm[j - 1] : 0.0);
        f[(size_t)j * a.PP] = make_float2((float)l, (float)(1.0 / piv));
    }
}

// Launch bounds: at most 256 registers per lane for both plane heights, so
// the products accumulate in VGPRs.  The PP = 64 build with a 512-register
// budget kept accumulators in AGPRs and its second step came out wrong (1e-6
// off the kernel-per-stage result, at -O3 only, while -O1 matched it to
// 1e-15); the VGPR-only build matches to rounding.
template <int PP, int JP>
__global__ __launch_bounds__(PP * 4, 128 / PP) void k_refine_plane(ModeArgs a, const Fields F, const double *bn,
                                                           const float *wf32, const float *wi32, const float *fac,
                                                           int steps, int m64, int nplanes)
{
    typedef Mfma<float> MM;
    typedef typename MM::acc acc_t;
    constexpr int NT = PP * 4;    // threads: PP / 16 wavefronts
    constexpr int U = JP / 16;    // 16-column MFMA tiles per wavefront
    constexpr int KS = PP / 4;    // k-steps of the 16x16x4 product
    constexpr int R = JP / 64;    // r rows per lane
    constexpr int LP = JP + 17;   // LDS row pitch (floats)
    constexpr int CH = 32;        // rows per chunk of the first forward product
    constexpr int CPER = CH * JP / NT;  // chunk elements per thread
    static_assert(CPER * NT == CH * JP, "chunk split");
    __shared__ float S[PP * LP];
    __shared__ float cS[JP];  // CM (zero past M)

    const int tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const int NZ = a.NZ, M = a.M;

    // products over the plane in LDS: acc[u] += sum_ks A[16w+lr][4ks+lk] B[4ks+lk][16u+lr]
    // for ks0 <= ks < ks1.  A comes from the binary32 weight table W32
    // ([red][row], PP x PP, zero past NZ: rows p >= NZ of Y and k >= NZ of X
    // come out zero) one k-step ahead, the B operands of the next k-step are
    // read from LDS while the current one is multiplied.
    auto product = [&](const float *W32, acc_t (&acc)[U], int ks0, int ks1) {
        const float *wp = W32 + (size_t)lk * PP + 16 * w + lr;
        float an = wp[(size_t)(4 * ks0) * PP];
        float bn_[U];
#pragma unroll
        for (int u = 0; u < U; ++u) bn_[u] = S[(4 * ks0 + lk) * LP + 16 * u + lr];
#pragma unroll 1
        for (int ks = ks0; ks < ks1; ++ks) {
            const float ac = an;
            float bc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) bc[u] = bn_[u];
            if (ks + 1 < ks1) {
                an = wp[(size_t)(4 * (ks + 1)) * PP];
#pragma unroll
                for (int u = 0; u < U; ++u) bn_[u] = S[(4 * (ks + 1) + lk) * LP + 16 * u + lr];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u] = MM::mac(ac, bc[u], acc[u]);
        }
    };
    // accumulators -> LDS rows 16w + 4lk + v, columns 16u + lr
    auto put = [&](const acc_t (&acc)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) S[(16 * w + MM::row(lane, v)) * LP + 16 * u + lr] = acc[u][v];
    };
    auto rsolve = [&](int m) {
        if (m <= m64) plane_rsolve<R, LP, double, 2>(a, S, m, w, lane);
        else if (w < PP / 64) plane_thomas<LP, JP>(S, cS, fac + (size_t)2 * m * PP * JP, a.PP, NZ, w, lane);
    };

    {
        // XCD-aware plane order: the blocks of one XCD (block b runs on XCD
        // b mod 8) take a contiguous run of planes sorted by order m, fields
        // fastest, so the planes sharing m's r factors meet in one L2
        const int nf = nplanes / a.L;
        const int chunk = (nplanes + 7) / 8;
        const int sidx = (blockIdx.x & 7) * chunk + (blockIdx.x >> 3);
        if (sidx >= nplanes) return;  // the whole block leaves together
        const int qs = sidx / nf, fi = sidx % nf;
        // qs: 0 -> m 0; 2m-1 -> cosine of m; 2m -> sine of m (m < N); L-1 -> m N
        const int q = qs == 0 ? 0 : (qs == a.L - 1 ? a.N : ((qs & 1) ? (qs + 1) / 2 : a.N + qs / 2));
        const int m = q <= a.N ? q : q - a.N;
        double *spec = static_cast<double *>(F.spec[fi]) + (size_t)q * a.K * a.J;
        for (int j = tid; j < JP; j += NT) cS[j] = (float)a.cm[j];  // read after the barriers below

        // 1. z forward of B, chunk by chunk
        acc_t acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = MM::zero();
        double cb[CPER];
        auto fetch = [&](int c) {
#pragma unroll
            for (int i = 0; i < CPER; ++i) {
                const int e = tid + NT * i, k = c * CH + e / JP, j = e % JP;
                const int kc = k < NZ ? k : NZ - 1, jc = j < M ? j : M - 1;
                cb[i] = spec[(size_t)kc * a.J + jc];
            }
        };
        auto stash = [&](int c) {
#pragma unroll
            for (int i = 0; i < CPER; ++i) {
                const int e = tid + NT * i, k = c * CH + e / JP, j = e % JP;
                S[k * LP + j] = (k < NZ && j < M) ? (float)cb[i] : 0.0f;
            }
        };
        fetch(0);
        stash(0);
        __syncthreads();
        constexpr int NCH = PP / CH;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if (c + 1 < NCH) fetch(c + 1);
            product(wf32, acc, c * (CH / 4), (c + 1) * (CH / 4));
            if (c + 1 < NCH) stash(c + 1);
            __syncthreads();
        }
        put(acc);
        __syncthreads();

        // 2. r solves
        rsolve(m);
        __syncthreads();

        // 3. z inverse
        acc_t x1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x1[u] = MM::zero();
        product(wi32, x1, 0, KS);

        if (steps == 1) {
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = 16 * w + MM::row(lane, v), j = 16 * u + lr;
                    if (k < NZ && j < M) spec[(size_t)k * a.J + j] = (double)x1[u][v];
                }
            return;
        }

        // X1 to the binary32 scratch plane (read back for the final sum) and
        // to LDS (the residual's stencil)
        float *xs = static_cast<float *>(F.aux[fi]) + (size_t)q * PP * JP;
        __syncthreads();  // every wavefront has read Y
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * w + MM::row(lane, v), j = 16 * u + lr;
                S[r * LP + j] = x1[u][v];
                xs[(size_t)r * JP + j] = x1[u][v];
            }
        __syncthreads();

        // 4. residual of the plane's system at X1, binary64, rounded to
        // binary32; cells past NZ or M are zero
        {
            float rr[U][4];
            double an4[4], bd4[4], cn4[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = 16 * w + MM::row(lane, v), kc = k < NZ ? k : NZ - 1;
                an4[v] = a.an_dev[kc];
                bd4[v] = bn[kc];
                cn4[v] = a.cn_dev[kc];
            }
            // B of UG column groups at a time, the next batch loaded while
            // the current one is used
            constexpr int UG = U < 2 ? U : 2;
            double bcur[UG][4], bnxt[UG][4];
            auto load_b = [&](double (&bb)[UG][4], int u0) {
#pragma unroll
                for (int g = 0; g < UG; ++g) {
                    const int j = 16 * (u0 + g) + lr, jc = j < M ? j : M - 1;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int k = 16 * w + MM::row(lane, v), kc = k < NZ ? k : NZ - 1;
                        bb[g][v] = spec[(size_t)kc * a.J + jc];
                    }
                }
            };
            load_b(bcur, 0);
#pragma unroll
            for (int u0 = 0; u0 < U; u0 += UG) {
                if (u0 + UG < U) load_b(bnxt, u0 + UG);
#pragma unroll
                for (int g = 0; g < UG; ++g) {
                    const int u = u0 + g;
                    const int j = 16 * u + lr, jc = j < M ? j : M - 1;
                    const double amj = a.am[jc], cmj = a.cm[jc], bmj = a.bm[(size_t)m * a.JP + jc];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        // every read in bounds (clamped), the out-of-system
                        // neighbours weighted by zero: no branches
                        const int k = 16 * w + MM::row(lane, v);
                        const int km = k > 0 ? k - 1 : k, kp = k + 1 < PP ? k + 1 : k;
                        const int jm = j > 0 ? j - 1 : j;
                        const float sx = S[k * LP + j], sm = S[km * LP + j], sp = S[kp * LP + j];
                        const float sl = S[k * LP + jm], sr = S[k * LP + j + 1];
                        const double xm = k > 0 ? (double)sm : 0.0;
                        const double xp = k + 1 < NZ ? (double)sp : 0.0;
                        const double xl = j > 0 ? (double)sl : 0.0;
                        const double xr = j + 1 < M ? (double)sr : 0.0;
                        const double ax =
                            an4[v] * xm + (bd4[v] + bmj) * (double)sx + cn4[v] * xp + amj * xl + cmj * xr;
                        rr[u][v] = (k < NZ && j < M) ? (float)(bcur[g][v] - ax) : 0.0f;
                    }
                }
                // one batch of B loads in flight at a time
                __builtin_amdgcn_sched_barrier(0);
                if (u0 + UG < U) {
#pragma unroll
                    for (int g = 0; g < UG; ++g)
#pragma unroll
                        for (int v = 0; v < 4; ++v) bcur[g][v] = bnxt[g][v];
                }
            }
            __syncthreads();  // every wavefront has read X1's stencil
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) S[(16 * w + MM::row(lane, v)) * LP + 16 * u + lr] = rr[u][v];
            __syncthreads();
        }

        // 5. the correction: z forward of R, r solves, z inverse
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = MM::zero();
        product(wf32, acc, 0, KS);
        __syncthreads();
        put(acc);
        __syncthreads();
        rsolve(m);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = MM::zero();
        product(wi32, acc, 0, KS);

        // 6. X = X1 + C in binary64 into the spectrum (this wavefront wrote
        // the X1 values it reads back)
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = 16 * w + MM::row(lane, v), j = 16 * u + lr;
                if (k < NZ && j < M)
                    spec[(size_t)k * a.J + j] = (double)xs[(size_t)k * JP + j] + (double)acc[u][v];
            }
    }
}

template <int PP, int JP>
void launch_refine_plane(const chy_grid *g, const ModeArgs &a, const Fields &F, int steps, int m64, hipStream_t s)
{
    const int nplanes = g->L * F.nf;
    const int blocks = 8 * ((nplanes + 7) / 8);
    hipLaunchKernelGGL((k_refine_plane<PP, JP>), dim3(blocks), dim3(PP * 4), 0, s, a, F, g->d.rz_bn, g->d.wf32,
                       g->d.wi32, g->d.fac32, steps, m64, nplanes);
}

}  // namespace

int launch_spec_residual_f(const chy_grid *g, const Fields &F, bool x32, hipStream_t s)
{
    ModeArgs a = make_args(g);
    a.an_dev = g->d.rz_an;
    a.cn_dev = g->d.rz_cn;
    const dim3 grid((unsigned)(((g->NZ + RROWS - 1) / RROWS) * ((g->M + 255) / 256)), g->L, F.nf);
    if (x32)
        hipLaunchKernelGGL(k_spec_residual<float>, grid, dim3(256), 0, s, a, F, g->d.rz_bn);
    else
        hipLaunchKernelGGL(k_spec_residual<double>, grid, dim3(256), 0, s, a, F, g->d.rz_bn);
    return (int)hipGetLastError();
}

int launch_thomas_pivots(const chy_grid *g, hipStream_t s)
{
    ModeArgs a = make_args(g);
    const int n = (g->N + 1) * g->NZ;
    hipLaunchKernelGGL(k_thomas_pivots, dim3((n + 255) / 256), dim3(256), 0, s, a, g->d.rinv);
    return (int)hipGetLastError();
}

int launch_thomas_factors(const chy_grid *g, hipStream_t s)
{
    const ModeArgs a = make_args(g);
    const int n = (g->N + 1) * g->NZ;
    hipLaunchKernelGGL(k_thomas_factors, dim3((n + 255) / 256), dim3(256), 0, s, a, g->d.fac32);
    return (int)hipGetLastError();
}

bool refine_planes_ok(const chy_grid *g, int steps)
{
    return (steps == 1 || steps == 2) && (g->PP == 64 || g->PP == 128) && g->JP <= 256 && g->NZ >= 2;
}

int launch_refine_planes_f(const chy_grid *g, const Fields &F, int steps, int m_fp64, hipStream_t s)
{
    if (!refine_planes_ok(g, steps)) return (int)hipErrorInvalidValue;
    ModeArgs a = make_args(g);
    a.an_dev = g->d.rz_an;
    a.cn_dev = g->d.rz_cn;
    const int m64 = std::max(-1, std::min(m_fp64, g->N));
    if (g->PP == 128) {
        switch (g->JP) {
        case 64: launch_refine_plane<128, 64>(g, a, F, steps, m64, s); break;
        case 128: launch_refine_plane<128, 128>(g, a, F, steps, m64, s); break;
        default: launch_refine_plane<128, 256>(g, a, F, steps, m64, s); break;
        }
    } else {
        switch (g->JP) {
        case 64: launch_refine_plane<64, 64>(g, a, F, steps, m64, s); break;
        case 128: launch_refine_plane<64, 128>(g, a, F, steps, m64, s); break;
        default: launch_refine_plane<64, 256>(g, a, F, steps, m64, s); break;
        }
    }
    return (int)hipGetLastError();
}


}  // namespace chy
