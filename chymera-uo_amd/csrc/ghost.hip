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

// argmax of |v| with the smallest index winning ties, across one block
__device__ void block_argmax_abs(double v, int idx, double *sv, int *si, double &out_v, int &out_i)
{
    const int tid = threadIdx.x;
    sv[tid] = v;
    si[tid] = idx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const double a = sv[tid], b = sv[tid + s];
            const int ia = si[tid], ib = si[tid + s];
            const bool takeb = (fabs(b) > fabs(a)) || (fabs(b) == fabs(a) && ib < ia);
            if (takeb) { sv[tid] = b; si[tid] = ib; }
        }
        __syncthreads();
    }
    out_v = sv[0];
    out_i = si[0];
    __syncthreads();
}

template <int NP>
__global__ __launch_bounds__(256) void k_zaxphi(GhostArgs a)
{
    __shared__ double sv[256];
    __shared__ int si[256];
    const int K = blockIdx.x + 2;  // Fortran K
    const int tid = threadIdx.x;
    const int N = a.L / 2;
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t row = (size_t)a.JS * (K - 1);
    double *phi = a.phi;

    // ZAXPHI abscissae (pot3.f:336-348)
    constexpr int np = NP;
    constexpr int ichk = np % 2, nn = np - ichk, num = nn / 2, ispecl = num + num + 1, jsp = num + 1;
    double x[11];
#pragma unroll
    for (int jj = 2; jj <= jsp; ++jj) {
        x[num + (jj - 1)] = a.zx[jj - 1];
        x[num - (jj - 2)] = -a.zx[jj - 1];
    }
    if (ichk == 1) x[ispecl] = a.zx[num + 1];

    // extrapolated value PH(L) for L=1..N (pot3.f:354-379)
    double best = 0.0;
    int besti = 0x7fffffff;
    for (int l = tid; l < N; l += blockDim.x) {
        double p2[11];
#pragma unroll
        for (int jj = 2; jj <= jsp; ++jj) {
            p2[num + (jj - 1)] = phi[row + (jj - 1) + plane * l];
            p2[num - (jj - 2)] = phi[row + (jj - 1) + plane * (l + N)];
        }
        if (ichk == 1) p2[ispecl] = phi[row + (num + 1) + plane * l];
#pragma unroll
        for (int i = 1; i <= np - 1; ++i) {
            const double p1 = p2[i], xinv = 1.0 / x[i], xr = 0.0 * xinv;
#pragma unroll
            for (int jj = i + 1; jj <= np; ++jj) {
                const double xratio = x[jj] * xinv, pb = p2[jj];
                p2[jj] = (p1 * (xratio - xr) + pb * (xr - 1.0)) / (xratio - 1.0);
            }
        }
        const double ph = p2[np];
        phi[row + (a.J + 1) + plane * l] = ph;
        phi[row + (a.J + 1) + plane * (l + N)] = ph;
        if (fabs(ph) > fabs(best) || (fabs(ph) == fabs(best) && l < besti)) { best = ph; besti = l; }
    }
    double phmax;
    int imax;
    block_argmax_abs(best, besti, sv, si, phmax, imax);
    // largest |1 - PH/PHMAX| (pot3.f:381-386); PH re-read from what this block wrote
    __threadfence_block();
    double err = 0.0;
    int erri = 0x7fffffff;
    for (int l = tid; l < N; l += blockDim.x) {
        const double ph = phi[row + (a.J + 1) + plane * l];
        const double er = 1.0 - ph / phmax;
        if (fabs(er) > fabs(err) || (fabs(er) == fabs(err) && fabs(er) > 0.0 && l < erri)) { err = er; erri = l; }
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
    case 1: hipLaunchKernelGGL(k_zaxphi<1>, dim3(g->K), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_zaxphi<2>, dim3(g->K), dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_zaxphi<3>, dim3(g->K), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_zaxphi<4>, dim3(g->K), dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL(k_zaxphi<5>, dim3(g->K), dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL(k_zaxphi<6>, dim3(g->K), dim3(256), 0, s, a); break;
    case 7: hipLaunchKernelGGL(k_zaxphi<7>, dim3(g->K), dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_zaxphi<8>, dim3(g->K), dim3(256), 0, s, a); break;
    case 9: hipLaunchKernelGGL(k_zaxphi<9>, dim3(g->K), dim3(256), 0, s, a); break;
    case 10: hipLaunchKernelGGL(k_zaxphi<10>, dim3(g->K), dim3(256), 0, s, a); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

}  // namespace chy
