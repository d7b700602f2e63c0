// ZAXPHI (pot/pot3.f:311-399) after the inverse FFT has written PHI on
// J=1..JMAX+1, K=1..KMAX+1 (its ghost cells included, fft.hip).
//   k_zaxphi  one block per K=2..KMAX+1: the NPOINT Aitken extrapolation of
//             PHI to r=0 from x=+-RHF(2..) written to PHI(JMAX+2,K,L) and
//             PHI(JMAX+2,K,L+N), and PHAC(K), the largest azimuthal spread of
//             that value (first index wins ties, as the reference's
//             sequential strict-greater scans do).  The last block to finish
//             (a ticket counter) takes PHICHK/KLOCAT over all PHAC: the PHAC
//             of largest magnitude, first K wins, KLOCAT unchanged when every
//             PHAC is 0 (pot3.f:391-399); it resets the counter for the next
//             launch.
#include "chypot_impl.h"

namespace chy {
namespace {

struct GhostArgs {
    double *phi;
    const double *zx;  // RHF(1..11)
    double *dscal;     // [0]=tmass [1]=phichk [2]=klocat [4+K]=PHAC(K)
    int J, K, L, JS, KS, npoint;
    int klocat_in;     // KLOCAT before the call (kept when every PHAC is 0)
    unsigned *ticket;  // blocks finished (0 between launches)
};

// (v, i) ordered by |v| first, then the smaller index: the reference's
// sequential strict-greater scans keep the first of equal magnitudes.  A total
// order, so any reduction tree gives the scan's answer.
__device__ __forceinline__ void take_larger(double &v, int &i, double v2, int i2)
{
    if (fabs(v2) > fabs(v) || (fabs(v2) == fabs(v) && i2 < i)) { v = v2; i = i2; }
}

// argmax over the block (blockDim.x a multiple of 64, at most 256): wavefront
// shuffles, then the wavefront winners through LDS.  Every thread gets the result.
__device__ void block_argmax_abs(double v, int idx, double *sv, int *si, double &out_v, int &out_i)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        take_larger(v, idx, __shfl_xor(v, off, 64), __shfl_xor(idx, off, 64));
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) { sv[w] = v; si[w] = idx; }
    __syncthreads();
    out_v = sv[0];
    out_i = si[0];
    for (int x = 1; x < nw; ++x) take_larger(out_v, out_i, sv[x], si[x]);
    __syncthreads();  // sv/si free for the next call
}

// PPT: values of L (1..LMAX/2) per thread, blockDim.x * PPT >= LMAX/2
template <int NP, int PPT>
__global__ __launch_bounds__(256) void k_zaxphi(GhostArgs a)
{
    __shared__ double sv[4];
    __shared__ int si[4];
    const int K = blockIdx.x + 2;  // Fortran K
    const int tid = threadIdx.x;
    const int N = a.L / 2;
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t row = (size_t)a.JS * (K - 1);
    double *phi = a.phi;

    // ZAXPHI abscissae (pot3.f:336-348)
    constexpr int np = NP;
    constexpr int ichk = np % 2, nn = np - ichk, num = nn / 2, ispecl = num + num + 1, jsp = num + 1;
    // every load first: the samples of this thread's L values and the abscissae
    double p2[PPT][11];
#pragma unroll
    for (int t = 0; t < PPT; ++t) {
        const int l = tid + (int)blockDim.x * t;
        const size_t lc = l < N ? (size_t)l : 0;
#pragma unroll
        for (int jj = 2; jj <= jsp; ++jj) {
            p2[t][num + (jj - 1)] = phi[row + (jj - 1) + plane * lc];
            p2[t][num - (jj - 2)] = phi[row + (jj - 1) + plane * (lc + N)];
        }
        if (ichk == 1) p2[t][ispecl] = phi[row + (num + 1) + plane * lc];
    }
    double x[11];
#pragma unroll
    for (int jj = 2; jj <= jsp; ++jj) {
        x[num + (jj - 1)] = a.zx[jj - 1];
        x[num - (jj - 2)] = -a.zx[jj - 1];
    }
    if (ichk == 1) x[ispecl] = a.zx[num + 1];

    // extrapolated value PH(L) (pot3.f:354-379), same operations in the same order
    double ph[PPT];
    double best = 0.0;
    int besti = 0x7fffffff;
#pragma unroll
    for (int t = 0; t < PPT; ++t) {
        const int l = tid + (int)blockDim.x * t;
#pragma unroll
        for (int i = 1; i <= np - 1; ++i) {
            const double p1 = p2[t][i], xinv = 1.0 / x[i], xr = 0.0 * xinv;
#pragma unroll
            for (int jj = i + 1; jj <= np; ++jj) {
                const double xratio = x[jj] * xinv, pb = p2[t][jj];
                p2[t][jj] = (p1 * (xratio - xr) + pb * (xr - 1.0)) / (xratio - 1.0);
            }
        }
        ph[t] = p2[t][np];
        if (l < N) {
            phi[row + (a.J + 1) + plane * l] = ph[t];
            phi[row + (a.J + 1) + plane * (l + N)] = ph[t];
            take_larger(best, besti, ph[t], l);
        }
    }
    double phmax;
    int imax;
    block_argmax_abs(best, besti, sv, si, phmax, imax);
    // largest |1 - PH/PHMAX| (pot3.f:381-386); a zero deviation never sets
    // the index, as the strict comparison against ERR=0 never takes one
    double err = 0.0;
    int erri = 0x7fffffff;
#pragma unroll
    for (int t = 0; t < PPT; ++t) {
        const int l = tid + (int)blockDim.x * t;
        const double er = 1.0 - ph[t] / phmax;
        if (l < N && (fabs(er) > fabs(err) || (fabs(er) == fabs(err) && fabs(er) > 0.0 && l < erri))) {
            err = er;
            erri = l;
        }
    }
    double e;
    int ei;
    block_argmax_abs(err, erri, sv, si, e, ei);

    // PHICHK/KLOCAT by the last block to finish
    __shared__ bool last;
    if (tid == 0) {
        a.dscal[4 + K] = e;
        __threadfence();
        last = atomicAdd(a.ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double best2 = 0.0;
    int bk = 0x7fffffff;
    for (int kk = 2 + tid; kk <= a.K + 1; kk += blockDim.x) {
        const double v = __hip_atomic_load(&a.dscal[4 + kk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (fabs(v) > fabs(best2)) { best2 = v; bk = kk; }
    }
    double pc;
    int pk;
    block_argmax_abs(best2, bk, sv, si, pc, pk);
    if (tid == 0) {
        a.dscal[1] = pc;
        a.dscal[2] = fabs(pc) > 0.0 ? (double)pk : (double)a.klocat_in;
        *a.ticket = 0u;
    }
}

template <int NP>
int launch_np(const chy_grid *g, const GhostArgs &a, hipStream_t s)
{
    // one thread per L up to 256 threads (a whole number of wavefronts)
    const int N = g->L / 2;
    const int threads = N <= 64 ? 64 : (N <= 256 ? N : 256);
    const int ppt = (N + threads - 1) / threads;  // 1 or 2 (LMAX <= 1024)
    if (ppt == 1) hipLaunchKernelGGL((k_zaxphi<NP, 1>), dim3(g->K), dim3(threads), 0, s, a);
    else if (ppt == 2) hipLaunchKernelGGL((k_zaxphi<NP, 2>), dim3(g->K), dim3(threads), 0, s, a);
    else return (int)hipErrorInvalidValue;
    return (int)hipGetLastError();
}

}  // namespace

int launch_zaxphi(const chy_grid *g, int npoint, int klocat_in, double *phi, hipStream_t s)
{
    GhostArgs a{};
    a.klocat_in = klocat_in;
    a.phi = phi;
    a.zx = g->d.zx;
    a.dscal = g->d.dscal;
    a.ticket = g->d.ticket;
    a.J = g->J; a.K = g->K; a.L = g->L; a.JS = g->J + 2; a.KS = g->K + 2;
    a.npoint = npoint;
    switch (npoint) {  // 1..10 (clamped by the callers)
    case 1: return launch_np<1>(g, a, s);
    case 2: return launch_np<2>(g, a, s);
    case 3: return launch_np<3>(g, a, s);
    case 4: return launch_np<4>(g, a, s);
    case 5: return launch_np<5>(g, a, s);
    case 6: return launch_np<6>(g, a, s);
    case 7: return launch_np<7>(g, a, s);
    case 8: return launch_np<8>(g, a, s);
    case 9: return launch_np<9>(g, a, s);
    case 10: return launch_np<10>(g, a, s);
    default: return (int)hipErrorInvalidValue;
    }
}

}  // namespace chy
