// The disk's gravitational pull on the central star: find_force
// (pot/wiggle.F:99-135), which wiggle (pot/wiggle.F:15-93) calls every hydro
// step (radhydro/3dhyd-main.f:479-486) to advance the star by leapfrog.
// Every cell is a point mass RHO*RHF*VOL at its centre (VOL = DR*DZ*DPHI*2,
// both halves of the disk) and pulls the star at (x, y, 0) with softening
// GSOFT:
//
//   (fx, fy) = sum over L=1..LMAX, K=2..KMAX+1, J=JMIN-3..JMAX+1 of
//              RHO*RHF*VOL * (xx, yy) / sqrt(xx^2 + yy^2 + ZHF^2 + GSOFT^2)^3,
//   xx = RHF cos(phi_L) - x,  yy = RHF sin(phi_L) - y,  phi_L = L DPHI - DPHI/2.
//
// The reference sums in OpenMP reduction order (it depends on the thread
// count); here each thread sums its cells over a run of FORCE_LC planes, the
// block reduces its threads in a fixed tree, and one block adds the block
// partials in block order -- a fixed order, so results repeat run to run.
// (A last-block ticket instead of the second launch was slower: every one of
// the ~2000 blocks then fences and increments the same counter.)
//
// RHO is read with the leading dimensions (JMAX+2, KMAX+2) that find_force
// declares (pot/wiggle.F:103), as the reference reads the array it is passed.
// One thread per (J, K) and run of planes, over the flattened (J, K) range of
// the sum with J fastest: a wavefront reads 64 consecutive cells of a plane,
// and no block is left with a few live lanes when the row of NJ = JMAX-JMIN+5
// cells is not a multiple of 64.  cos/sin of phi_L come from a table.
#include <mutex>
#include <vector>

#include "chypot_impl.h"

namespace chy {
namespace {

struct ForceArgs {
    const double *rhf, *zhf, *cs, *rho;  // cs: (cos phi_L, sin phi_L) pairs
    double *part;                        // 2 doubles per block
    int j2, k2, j0, nj, kmax, lmax;      // RHO pitches JMAX+2, KMAX+2; first J (JMIN-3) and count
    double x, y, vol, gs2;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// the block's (fx, fy) to thread 0: wavefront sums, then the four wavefronts
// in a fixed tree
__device__ __forceinline__ void block_sum(double &fx, double &fy, double (&red)[2][4])
{
    const int t = (int)threadIdx.x;
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    if ((t & 63) == 0) {
        red[0][t >> 6] = fx;
        red[1][t >> 6] = fy;
    }
    __syncthreads();
    fx = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    fy = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
}

constexpr int FORCE_LC = 16;  // planes L per thread

__global__ __launch_bounds__(256) void k_find_force(ForceArgs a)
{
#pragma clang fp contract(off)
    const int t = (int)threadIdx.x;
    const int e = (int)blockIdx.x * 256 + t;  // flat (J, K) index of the sum, J fastest
    const int L0 = (int)blockIdx.y * FORCE_LC, L1 = min(L0 + FORCE_LC, a.lmax);
    double fx = 0.0, fy = 0.0;
    if (e < a.nj * a.kmax) {
        // each cell's terms in the order of wiggle.F:119-126
        const int kk = e / a.nj;
        const int J = a.j0 + (e - kk * a.nj), K = 2 + kk;
        const double rh = a.rhf[J - 1], zz = a.zhf[K - 1];
        const double z2 = zz * zz;
        const size_t plane = (size_t)a.j2 * a.k2;
        const double *rp = a.rho + (size_t)(J - 1) + (size_t)a.j2 * (K - 1);
        for (int L = L0; L < L1; ++L) {
            const double xx = rh * a.cs[2 * L] - a.x, yy = rh * a.cs[2 * L + 1] - a.y;
            const double s = sqrt(xx * xx + yy * yy + z2 + a.gs2);
            const double dd = 1.0 / (s * s * s);
            const double tm = rp[plane * L] * rh * a.vol;
            fx += tm * dd * xx;
            fy += tm * dd * yy;
        }
    }
    __shared__ double red[2][4];
    block_sum(fx, fy, red);
    if (t == 0) {
        const size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        a.part[2 * b] = fx;
        a.part[2 * b + 1] = fy;
    }
}

// One block adds the partials: each thread a strided share in order, then the
// same fixed tree.
__global__ __launch_bounds__(256) void k_force_total(const double *part, int nparts, double *out)
{
    const int t = (int)threadIdx.x;
    double fx = 0.0, fy = 0.0;
    for (int i = t; i < nparts; i += 256) {
        fx += part[2 * i];
        fy += part[2 * i + 1];
    }
    __shared__ double red[2][4];
    block_sum(fx, fy, red);
    if (t == 0) {
        out[0] = fx;
        out[1] = fy;
    }
}

// cos/sin of phi_L = L DPHI - DPHI/2 (the angle find_force forms,
// pot/wiggle.F:113) for the no-allocation form, which has no host table
__global__ __launch_bounds__(256) void k_force_trig(double *cs, int lmax, double dphi)
{
    const int l = (int)(blockIdx.x * 256 + threadIdx.x);
    if (l >= lmax) return;
    const double a = (double)(l + 1) * dphi - 0.5 * dphi;
    cs[2 * l] = cos(a);
    cs[2 * l + 1] = sin(a);
}

// Scratch of chy_find_force_noalloc: the block partials (two doubles each)
// followed by the cos/sin table, in device memory the library never
// allocates at run time.  One copy per GPU, shared by every call on it.
__device__ __align__(256) double scratch[2 * kForceNoAllocBlocks + 2 * kForceNoAllocL];

}  // namespace

// The scratch array's device address on `device` (host code may not use the
// array's name as a device pointer).  Taken once per device, with the device
// current, at the first context created on it.
int force_scratch_address(int device, double **out)
{
    static std::mutex mu;
    static std::vector<double *> addr;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return (int)hipErrorInvalidDevice;
    if ((int)addr.size() <= device) addr.resize((size_t)device + 1, nullptr);
    if (!addr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        addr[device] = static_cast<double *>(p);
    }
    *out = addr[device];
    return (int)hipSuccess;
}

int launch_find_force_noalloc(const chy_force_desc *d, const double *rhf, const double *zhf, const double *rho,
                              double x, double y, double *scr, double *out, hipStream_t st)
{
    double *part = scr, *cs = scr + 2 * (size_t)kForceNoAllocBlocks;
    hipLaunchKernelGGL(k_force_trig, dim3((unsigned)((d->lmax + 255) / 256)), dim3(256), 0, st, cs, d->lmax,
                       d->dphi);
    int rc = (int)hipGetLastError();
    if (rc) return rc;
    return launch_find_force(d, rhf, zhf, cs, rho, x, y, part, out, st);
}

static int force_row_blocks(const chy_force_desc *d)
{
    const int nj = d->jmax + 1 - (d->jmin - 3) + 1;
    return (nj * d->kmax + 255) / 256;
}

int force_blocks(const chy_force_desc *d)
{
    return force_row_blocks(d) * ((d->lmax + FORCE_LC - 1) / FORCE_LC);
}

int launch_find_force(const chy_force_desc *d, const double *rhf, const double *zhf, const double *cs,
                      const double *rho, double x, double y, double *part, double *out, hipStream_t st)
{
    ForceArgs a{};
    a.rhf = rhf; a.zhf = zhf; a.cs = cs; a.rho = rho; a.part = part;
    a.j2 = d->jmax + 2; a.k2 = d->kmax + 2;  // the leading dimensions wiggle.F:103 declares
    a.j0 = d->jmin - 3; a.nj = d->jmax + 1 - a.j0 + 1;
    a.kmax = d->kmax; a.lmax = d->lmax;
    a.x = x; a.y = y;
    a.vol = d->dr * d->dz * d->dphi * 2.0;
    a.gs2 = d->gsoft * d->gsoft;
    const dim3 grid(force_row_blocks(d), (d->lmax + FORCE_LC - 1) / FORCE_LC);
    hipLaunchKernelGGL(k_find_force, grid, dim3(256), 0, st, a);
    int rc = (int)hipGetLastError();
    if (rc) return rc;
    hipLaunchKernelGGL(k_force_total, dim3(1), dim3(256), 0, st, part, (int)(grid.x * grid.y), out);
    return (int)hipGetLastError();
}

}  // namespace chy
