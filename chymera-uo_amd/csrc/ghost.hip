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
#include "zax.h"

namespace chy {
namespace {

struct GhostArgs {
    double *phi;
    const double *zs;  // [KMAX][kZaxSamples][L] copy of PHI(J=2..6,K,L), or null: read PHI
    const double *zx;  // RHF(1..11)
    double *dscal;     // [0]=tmass [1]=phichk [2]=klocat [4+K]=PHAC(K)
    int J, K, L, JS, KS, npoint;
    int klocat_in;     // KLOCAT before the call (kept when every PHAC is 0)
    unsigned *ticket;  // blocks finished (0 between launches)
};

// PPT: values of L (1..LMAX/2) per thread, blockDim.x * PPT >= LMAX/2
template <int NP, int PPT>
__global__ __launch_bounds__(256) void k_zaxphi(GhostArgs a, const Fields F)
{
    // field f = blockIdx.y: its PHI, diagnostics and ticket
    a.phi = F.phi[blockIdx.y];
    if (a.zs) a.zs = F.zsamp[blockIdx.y];
    a.dscal = F.dscal[blockIdx.y];
    a.ticket = F.ticket[blockIdx.y];
    a.klocat_in = F.klocat_in[blockIdx.y];
    __shared__ double sv[4];
    __shared__ int si[4];
    const int K = blockIdx.x + 2;  // Fortran K
    const int tid = threadIdx.x;
    const int N = a.L / 2;
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t row = (size_t)a.JS * (K - 1);
    double *phi = a.phi;

    // every load first: the samples of this thread's L values, then the abscissae
    // PHI(J,K,L) and PHI(J,K,L+N) for J = 2..NUM+1, and PHI(NUM+2,K,L) for odd NP,
    // from the inverse FFT's contiguous copy when there is one (one 8-byte
    // piece per DRAM page otherwise)
    constexpr int num = (NP - NP % 2) / 2;
    double smp[PPT][2][num + 2];
    const double *zs = a.zs ? a.zs + (size_t)(K - 2) * kZaxSamples * a.L : nullptr;
#pragma unroll
    for (int t = 0; t < PPT; ++t) {
        const int l = tid + (int)blockDim.x * t;
        const size_t lc = l < N ? (size_t)l : 0;
        if (zs) {
#pragma unroll
            for (int jj = 2; jj <= num + 1; ++jj) {
                smp[t][0][jj - 1] = zs[(size_t)(jj - 2) * a.L + lc];
                smp[t][1][jj - 1] = zs[(size_t)(jj - 2) * a.L + lc + N];
            }
            if (NP % 2 == 1) smp[t][0][num + 1] = zs[(size_t)num * a.L + lc];
        } else {
#pragma unroll
            for (int jj = 2; jj <= num + 1; ++jj) {
                smp[t][0][jj - 1] = phi[row + (jj - 1) + plane * lc];
                smp[t][1][jj - 1] = phi[row + (jj - 1) + plane * (lc + N)];
            }
            if (NP % 2 == 1) smp[t][0][num + 1] = phi[row + (num + 1) + plane * lc];
        }
    }
    double x[11];
    zax_abscissae<NP>(a.zx, x);

    double ph[PPT];
#pragma unroll
    for (int t = 0; t < PPT; ++t) {
        const int l = tid + (int)blockDim.x * t;
        ph[t] = zax_extrapolate<NP>(x, [&](int jj, bool far) { return smp[t][far ? 1 : 0][jj - 1]; });
        if (l < N) {
            phi[row + (a.J + 1) + plane * l] = ph[t];
            phi[row + (a.J + 1) + plane * (l + N)] = ph[t];
        }
    }
    const double e = zax_phac<PPT>(ph, tid, (int)blockDim.x, N, sv, si);
    zax_finish(a.dscal, K, e, a.ticket, gridDim.x, a.K, a.klocat_in, sv, si);
}

template <int NP>
int launch_np(const chy_grid *g, const GhostArgs &a, const Fields &F, hipStream_t s)
{
    // one thread per L up to 256 threads (a whole number of wavefronts)
    const int N = g->L / 2;
    const int threads = N <= 64 ? 64 : (N <= 256 ? N : 256);
    const int ppt = (N + threads - 1) / threads;  // 1 or 2 (LMAX <= 1024)
    const dim3 grid(g->K, F.nf);
    if (ppt == 1) hipLaunchKernelGGL((k_zaxphi<NP, 1>), grid, dim3(threads), 0, s, a, F);
    else if (ppt == 2) hipLaunchKernelGGL((k_zaxphi<NP, 2>), grid, dim3(threads), 0, s, a, F);
    else return (int)hipErrorInvalidValue;
    return (int)hipGetLastError();
}

}  // namespace

bool zaxphi_from_samples(const chy_grid *g) { return g->J >= kZaxSamples; }

int launch_zaxphi_f(const chy_grid *g, const Fields &F, int npoint, bool samples, hipStream_t s)
{
    GhostArgs a{};
    if (samples) {
        if (!zaxphi_from_samples(g)) return (int)hipErrorInvalidValue;
        a.zs = F.zsamp[0];  // non-null marker: the kernel takes each field's
    }
    a.zx = g->d.zx;
    a.J = g->J; a.K = g->K; a.L = g->L; a.JS = g->J + 2; a.KS = g->K + 2;
    a.npoint = npoint;
    switch (npoint) {  // 1..10 (clamped by the callers)
    case 1: return launch_np<1>(g, a, F, s);
    case 2: return launch_np<2>(g, a, F, s);
    case 3: return launch_np<3>(g, a, F, s);
    case 4: return launch_np<4>(g, a, F, s);
    case 5: return launch_np<5>(g, a, F, s);
    case 6: return launch_np<6>(g, a, F, s);
    case 7: return launch_np<7>(g, a, F, s);
    case 8: return launch_np<8>(g, a, F, s);
    case 9: return launch_np<9>(g, a, F, s);
    case 10: return launch_np<10>(g, a, F, s);
    default: return (int)hipErrorInvalidValue;
    }
}

int launch_zaxphi(const chy_grid *g, int npoint, int klocat_in, double *phi, bool samples, hipStream_t s)
{
    const Fields F = one_field(g, nullptr, phi, nullptr, klocat_in);
    return launch_zaxphi_f(g, F, npoint, samples, s);
}

}  // namespace chy
