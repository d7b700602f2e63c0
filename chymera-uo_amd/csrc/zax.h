// The pieces of ZAXPHI (pot/pot3.f:311-399) that ghost.hip's k_zaxphi is
// built from: abscissae, the Aitken extrapolation, PHAC, and PHICHK/KLOCAT by
// the last block to finish.
#pragma once

#include <hip/hip_runtime.h>

namespace chy {

// (v, i) ordered by |v| first, then the smaller index: the reference's
// sequential strict-greater scans keep the first of equal magnitudes.  A total
// order, so any reduction tree gives the scan's answer.
__device__ __forceinline__ void take_larger(double &v, int &i, double v2, int i2)
{
    if (fabs(v2) > fabs(v) || (fabs(v2) == fabs(v) && i2 < i)) { v = v2; i = i2; }
}

// argmax over the block (blockDim.x a multiple of 64, at most 1024): wavefront
// shuffles, then the wavefront winners through LDS (sv, si: one entry per
// wavefront).  Every thread gets the result.
__device__ inline void block_argmax_abs(double v, int idx, double *sv, int *si, double &out_v, int &out_i)
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

// ZAXPHI's abscissae x(1..NP) from RHF(1..11) (pot3.f:336-348); zx is RHF.
template <int NP>
__device__ __forceinline__ void zax_abscissae(const double *zx, double (&x)[11])
{
    constexpr int ichk = NP % 2, nn = NP - ichk, num = nn / 2, ispecl = num + num + 1, jsp = num + 1;
#pragma unroll
    for (int jj = 2; jj <= jsp; ++jj) {
        x[num + (jj - 1)] = zx[jj - 1];
        x[num - (jj - 2)] = -zx[jj - 1];
    }
    if (ichk == 1) x[ispecl] = zx[num + 1];
}

// The NP-point Aitken extrapolation of PHI to r=0 (pot3.f:354-379), same
// operations in the same order.  at(J, far) gives PHI(J,K,L) (far=false) or
// PHI(J,K,L+LMAX/2) (far=true) for J = 2..NUM+2.
template <int NP, typename At>
__device__ __forceinline__ double zax_extrapolate(const double (&x)[11], At at)
{
    constexpr int ichk = NP % 2, nn = NP - ichk, num = nn / 2, ispecl = num + num + 1, jsp = num + 1;
    double p2[11];
#pragma unroll
    for (int jj = 2; jj <= jsp; ++jj) {
        p2[num + (jj - 1)] = at(jj, false);
        p2[num - (jj - 2)] = at(jj, true);
    }
    if (ichk == 1) p2[ispecl] = at(num + 2, false);
#pragma unroll
    for (int i = 1; i <= NP - 1; ++i) {
        const double p1 = p2[i], xinv = 1.0 / x[i], xr = 0.0 * xinv;
#pragma unroll
        for (int jj = i + 1; jj <= NP; ++jj) {
            const double xratio = x[jj] * xinv, pb = p2[jj];
            p2[jj] = (p1 * (xratio - xr) + pb * (xr - 1.0)) / (xratio - 1.0);
        }
    }
    return p2[NP];
}

// PHAC(K) from this block's extrapolated values ph[t] of L = l0 + stride t:
// PHMAX (largest |PH|, first L wins), then the deviation 1 - PH/PHMAX of
// largest magnitude, first L wins; a zero deviation never sets the index, as
// the strict comparison against ERR=0 never takes one (pot3.f:360-386).
template <int PPT>
__device__ __forceinline__ double zax_phac(const double (&ph)[PPT], int l0, int stride, int n, double *sv, int *si)
{
    double best = 0.0;
    int besti = 0x7fffffff;
#pragma unroll
    for (int t = 0; t < PPT; ++t) {
        const int l = l0 + stride * t;
        if (l < n) take_larger(best, besti, ph[t], l);
    }
    double phmax;
    int imax;
    block_argmax_abs(best, besti, sv, si, phmax, imax);
    double err = 0.0;
    int erri = 0x7fffffff;
#pragma unroll
    for (int t = 0; t < PPT; ++t) {
        const int l = l0 + stride * t;
        const double er = 1.0 - ph[t] / phmax;
        if (l < n && (fabs(er) > fabs(err) || (fabs(er) == fabs(err) && fabs(er) > 0.0 && l < erri))) {
            err = er;
            erri = l;
        }
    }
    double e;
    int ei;
    block_argmax_abs(err, erri, sv, si, e, ei);
    return e;
}

// PHICHK/KLOCAT (pot3.f:391-399) by the last of `nblocks` blocks to finish:
// each block has stored its PHAC(K) in dscal[4+K]; the one that takes the
// last ticket reduces them (largest |PHAC|, first K wins; KLOCAT unchanged
// when every PHAC is 0) and resets the ticket for the next launch.
__device__ inline void zax_finish(double *dscal, int K_phac, double phac, unsigned *ticket, unsigned nblocks,
                                  int kmax, int klocat_in, double *sv, int *si)
{
    __shared__ bool last;
    if (threadIdx.x == 0) {
        dscal[4 + K_phac] = phac;
        __threadfence();
        last = atomicAdd(ticket, 1u) == nblocks - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double best2 = 0.0;
    int bk = 0x7fffffff;
    for (int kk = 2 + (int)threadIdx.x; kk <= kmax + 1; kk += blockDim.x) {
        const double v = __hip_atomic_load(&dscal[4 + kk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (fabs(v) > fabs(best2)) { best2 = v; bk = kk; }
    }
    double pc;
    int pk;
    block_argmax_abs(best2, bk, sv, si, pc, pk);
    if (threadIdx.x == 0) {
        dscal[1] = pc;
        dscal[2] = fabs(pc) > 0.0 ? (double)pk : (double)klocat_in;
        *ticket = 0u;
    }
}

}  // namespace chy
