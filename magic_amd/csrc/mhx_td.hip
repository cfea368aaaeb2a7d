// lm-space assembly of the explicit terms of the time step at interior
// radial levels (src/get_td.f90): the step that follows the forward
// transforms of the nonlinear terms, for a rotating Boussinesq MHD run
// without the double-curl form:
//   get_dwdt (:133-203)  dwdt = or2 AdvrLM + Coriolis (dPhi dw, dTheta2 z)
//   get_dzdt (:379-444)  dzdt = dLh AdvpLM + Coriolis (dPhi z, dTheta3 dw,
//                        dTheta4 w)
//   get_dpdt (:314-377)  dpdt = -dLh AdvtLM + Coriolis (dPhi (dw + or1 dLh
//                        w), dTheta3 z); l = 0 left at zero
//   get_dsdt (:446-500)  dsdt = dLh VStLM, l = 0: epsc epscProf
//   get_dbdt (:530-589)  dbdt = dLh VxBpLM, djdt = dLh or4 VxBrLM,
//                        dVxBhLM = -dLh r^2 VxBtLM
// with the operators of src/horizontal.f90:201-229 and the neighbours l -/+ 1
// at the same m (src/blocking.f90:322-334); the Coriolis terms keep l < l_R,
// the l = l_R terms drop the l + 1 neighbour, and vanish above l_R.
// One thread per (lm, level), 256 consecutive lm per block: the l -/+ 1
// neighbours of a coefficient are the adjacent entries of its row (m-major,
// l fastest).  Memory bound: 10 complex in and 7 out per (lm, level).
// The same file holds the per-level kinetic-energy sums of the energy
// diagnostics (k_e_kin), another lm-space pass over the levels' fields.
#include "mhx_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace mhx {

namespace {

__device__ inline double2 cmulr(double a, double2 z) { return make_double2(a * z.x, a * z.y); }
__device__ inline double2 cadd2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub2(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// i m z
__device__ inline double2 cim(double m, double2 z) { return make_double2(-m * z.y, m * z.x); }

// sqrt((l+m)(l-m) / ((2l-1)(2l+1))), zero below l = m (src/horizontal.f90:201-205)
__device__ inline double clm(int l, int m)
{
    if (l <= m) return 0.0;
    const double dl = (double)l, dm = (double)m;
    return sqrt((dl + dm) * (dl - dm) / ((2.0 * dl - 1.0) * (2.0 * dl + 1.0)));
}

__global__ __launch_bounds__(256) void k_get_td(TdArgs a)
{
    const int lm = blockIdx.x * blockDim.x + threadIdx.x, lev = blockIdx.y;
    if (lm >= a.lm_max) return;
    const int l = a.lm2l[lm], m = a.lm2m[lm], L = a.l_max;
    const double dm = (double)m;
    const size_t i = (size_t)lev * a.ld_lm + (size_t)lm;
    const double or2 = a.rad[lev], or4 = a.rad[a.nlev + lev], epsc = a.rad[2 * a.nlev + lev];
    const int lR = a.lcut[lev];
    const double or1 = sqrt(or2), r2 = 1.0 / or2, c2 = 2.0 * a.CorFac;
    const double2 *in[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) in[k] = reinterpret_cast<const double2 *>(a.in[k]);
    const double2 *w = in[7], *dw = in[8], *z = in[9];
    const double2 zero = make_double2(0.0, 0.0);
    const double dl = (double)l, dLh = dl * (dl + 1.0);
    const double cS = clm(l, m), cA = clm(l + 1, m);
    const double t2S = (dl - 1.0) * cS, t2A = (dl + 2.0) * cA;
    const double t3S = (dl - 1.0) * (dl + 1.0) * cS, t3A = dl * (dl + 2.0) * cA;
    const double t4S = (dl + 1.0) * cS * ((dl - 1.0) * dl), t4A = dl * cA * ((dl + 1.0) * (dl + 2.0));
    // neighbours: l - 1 exists for l > m (its factors vanish at l = m),
    // l + 1 for l < l_max
    const size_t iS = l > m ? i - 1 : i;
    const bool hasA = l < L;
    const double2 zA = hasA ? z[i + 1] : zero, dwA = hasA ? dw[i + 1] : zero;
    const double2 wA = hasA ? w[i + 1] : zero;
    const double2 zS = z[iS], dwS = dw[iS], wS = w[iS];
    const double2 zl = z[i], dwl = dw[i], wl = w[i];
    double2 dwdt, dzdt, dpdt, dsdt;
    if (l == 0) {
        // lm = 1 of the reference (m = 0 only)
        dwdt = cadd2(cmulr(or2, in[0][i]), cmulr(c2 * or1 * t2A, zA));
        dzdt = cmulr(c2 * or2, cadd2(cmulr(t3A, dwA), cmulr(or1 * t4A, wA)));
        dpdt = zero;
        dsdt = make_double2(epsc, 0.0);
    } else {
        double2 cw = zero, cz = zero, cp = zero;
        if (l <= lR) {
            const bool below = l < lR;  // the l + 1 neighbour enters below l_R only
            cw = csub2(cim(dm, dwl), cmulr(t2S, zS));
            if (below) cw = cadd2(cw, cmulr(t2A, zA));
            cw = cmulr(c2 * or1, cw);
            cz = csub2(cadd2(cim(dm, zl), cmulr(t3S, dwS)), cmulr(or1 * t4S, wS));
            if (below) cz = cadd2(cz, cadd2(cmulr(t3A, dwA), cmulr(or1 * t4A, wA)));
            cz = cmulr(c2 * or2, cz);
            cp = cadd2(cmulr(-1.0, cim(dm, cadd2(dwl, cmulr(or1 * dLh, wl)))), cmulr(t3S, zS));
            if (below) cp = cadd2(cp, cmulr(t3A, zA));
            cp = cmulr(c2 * or2, cp);
        }
        dwdt = cadd2(cmulr(or2, in[0][i]), cw);
        dzdt = cadd2(cmulr(dLh, in[2][i]), cz);
        dpdt = cadd2(cmulr(-dLh, in[1][i]), cp);
        dsdt = cmulr(dLh, in[3][i]);
    }
    double2 *o[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) o[k] = reinterpret_cast<double2 *>(a.out[k]);
    o[0][i] = dwdt;
    o[1][i] = dzdt;
    o[2][i] = dpdt;
    o[3][i] = dsdt;
    o[4][i] = cmulr(dLh, in[6][i]);
    o[5][i] = cmulr(dLh * or4, in[4][i]);
    o[6][i] = cmulr(-dLh * r2, in[5][i]);
}

// Per-level sums of the kinetic energy (src/kinetic_energy.f90:119-160):
// e_p = orho1 dLh (dLh or2 |w|^2 + |dw|^2), e_t = orho1 dLh |z|^2 with
// |c|^2 = Re(c)^2 at m = 0 and 2 |c|^2 otherwise, split into all, m = 0
// (axisymmetric), l + m even / odd (equatorially symmetric poloidal /
// toroidal) and m = 0 with l even / odd.  Blocks (chunk, level): each
// thread accumulates a strided subset of the chunk's lm, a fixed tree over
// the block gives the chunk's partial sums, and k_e_kin_sum adds the
// partials of a level in chunk order, so the result does not depend on
// timing.
__global__ __launch_bounds__(256) void k_e_kin(EkArgs a)
{
    const int lev = blockIdx.y, nch = gridDim.x;
    const double2 *w = reinterpret_cast<const double2 *>(a.w) + (size_t)lev * a.ld_lm;
    const double2 *dw = reinterpret_cast<const double2 *>(a.dw) + (size_t)lev * a.ld_lm;
    const double2 *z = reinterpret_cast<const double2 *>(a.z) + (size_t)lev * a.ld_lm;
    const double or2 = a.or2[lev], orho = (a.mag || !a.orho1) ? 1.0 : a.orho1[lev];
    // the magnetic sums take the opposite l parity for their equatorially
    // symmetric parts (src/magnetic_energy.f90:281-300 against
    // src/kinetic_energy.f90:143-157)
    const int par = a.mag ? 1 : 0;
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int lm = blockIdx.x * blockDim.x + threadIdx.x; lm < a.lm_max; lm += nch * blockDim.x) {
        const int l = a.lm2l[lm], m = a.lm2m[lm];
        const double dLh = (double)(l * (l + 1));
        const double2 cw = w[lm], cd = dw[lm], cz = z[lm];
        double nw, nd, nz;
        if (m == 0) {
            nw = cw.x * cw.x;
            nd = cd.x * cd.x;
            nz = cz.x * cz.x;
        } else {
            nw = 2.0 * (cw.x * cw.x + cw.y * cw.y);
            nd = 2.0 * (cd.x * cd.x + cd.y * cd.y);
            nz = 2.0 * (cz.x * cz.x + cz.y * cz.y);
        }
        const double ep = orho * dLh * (dLh * or2 * nw + nd);
        const double et = orho * dLh * nz;
        acc[0] += ep;
        acc[1] += et;
        if (m == 0) {
            acc[2] += ep;
            acc[3] += et;
            if ((l & 1) == par) acc[6] += ep;
            else acc[7] += et;
        }
        if (((l + m) & 1) == par) acc[4] += ep;
        else acc[5] += et;
        if (l == 1) {
            acc[8] += ep;               // dipole
            if (m == 0) acc[9] += ep;   // axial dipole
        }
    }
    __shared__ double red[10][256];
    for (int k = 0; k < 10; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int k = 0; k < 10; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
        __syncthreads();
    }
    if ((int)threadIdx.x < a.nsum)
        a.part[((size_t)threadIdx.x * a.nlev + lev) * nch + blockIdx.x] = red[threadIdx.x][0];
}

__global__ void k_e_kin_sum(const double *__restrict__ part, int n, int nch, double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // k * nlev + level
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[(size_t)i * nch + c];
    out[i] = s;
}

}  // namespace

int launch_e_kin(const Plan *p, const EkArgs &args, hipStream_t s)
{
    if (args.nlev <= 0) return MHX_OK;
    EkArgs a = args;
    a.lm2l = p->d_lm2l;
    a.lm2m = p->d_lm2m;
    a.lm_max = p->lm_max;
    const int nch = std::min(EK_CHUNKS, (a.lm_max + 255) / 256);
    k_e_kin<<<dim3((unsigned)nch, (unsigned)a.nlev), 256, 0, s>>>(a);
    MHX_HIP_CHECK(hipGetLastError());
    const int n = a.nsum * a.nlev;
    k_e_kin_sum<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a.part, n, nch, a.out);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

int launch_get_td(const Plan *p, const TdArgs &args, hipStream_t s)
{
    if (args.nlev <= 0) return MHX_OK;
    TdArgs a = args;
    a.lm2l = p->d_lm2l;
    a.lm2m = p->d_lm2m;
    a.l_max = p->l_max;
    a.lm_max = p->lm_max;
    dim3 grid((unsigned)((p->lm_max + 255) / 256), (unsigned)a.nlev);
    k_get_td<<<grid, 256, 0, s>>>(a);
    MHX_HIP_CHECK(hipGetLastError());
    return MHX_OK;
}

}  // namespace mhx

// ---------------------------------------------------------------------------
// Allocation-free kinetic-energy sums (mhx_get_e_kin_noalloc): the kernels of
// mhx_get_e_kin on device arrays only, with their partial sums in a
// file-scope device array instead of plan memory that grows with hipMalloc.
// Each device holds its own copy of the array (until the process ends); its
// address on a device comes from hipGetSymbolAddress, looked up on the entry's
// first call on that device with the device current.  Every call on a device
// shares that device's array: calls that overlap on one device race.  No entry
// needs clearing: every partial sum the final reduction reads is written by
// the same call.

namespace mhx {
namespace {

constexpr int EK_NOALLOC_MAX_LEVELS = 2048;
constexpr int EK_NOALLOC_MAX_DEVICES = 64;

// the partial sums: 10 sums x levels x EK_CHUNKS chunks (2.6 MB)
static __device__ __align__(256) double scratch[10 * (size_t)EK_CHUNKS * EK_NOALLOC_MAX_LEVELS];

std::mutex g_scratch_mu;
double *g_scratch_ptr[EK_NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on device dev (current), looked up once
int scratch_on_device(int dev, double **out)
{
    if (dev < 0 || dev >= EK_NOALLOC_MAX_DEVICES)
        return fail(MHX_E_UNSUPPORTED, "device ordinal " + std::to_string(dev) + " beyond " +
                                           std::to_string(EK_NOALLOC_MAX_DEVICES));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(MHX_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MHX_OK;
}

}  // namespace

}  // namespace mhx

using namespace mhx;

int mhx_get_e_kin_noalloc(mhx_plan *h, int nlev, const double *w, const double *dw, const double *z,
                          size_t ld_lm, const double *or2, const double *orho1, double *sums,
                          void *stream)
{
    if (!h) return fail(MHX_E_INVALID, "null plan");
    Plan *P = &h->p;
    if (nlev <= 0) return MHX_OK;
    if (nlev > EK_NOALLOC_MAX_LEVELS)
        return fail(MHX_E_INVALID, "mhx_get_e_kin_noalloc: " + std::to_string(nlev) +
                                       " levels exceed its fixed scratch of " +
                                       std::to_string(EK_NOALLOC_MAX_LEVELS) + " levels");
    if (!w || !dw || !z || !or2 || !orho1 || !sums) return fail(MHX_E_INVALID, "null pointer");
    if (ld_lm < (size_t)P->lm_max) return fail(MHX_E_INVALID, "ld_lm below lm_max");
    const void *arrs[6] = {w, dw, z, or2, orho1, sums};
    for (const void *a : arrs)
        if (!is_device_ptr(a))
            return fail(MHX_E_INVALID, "mhx_get_e_kin_noalloc takes device arrays only (a host array "
                                       "would need staging memory)");
    int dev = -1;
    MHX_HIP_CHECK(hipGetDevice(&dev));
    if (dev != P->device)
        return fail(MHX_E_INVALID, "the current device is not the plan's device");
    double *part = nullptr;
    int rc = scratch_on_device(dev, &part);
    if (rc) return rc;
    EkArgs a;
    std::memset(&a, 0, sizeof(a));
    a.w = w;
    a.dw = dw;
    a.z = z;
    a.ld_lm = ld_lm;
    a.or2 = or2;
    a.orho1 = orho1;
    a.part = part;
    a.out = sums;
    a.nlev = nlev;
    a.mag = 0;
    a.nsum = 8;
    return launch_e_kin(P, a, (hipStream_t)stream);
}
