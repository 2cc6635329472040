// Delay tabs and apodisations of a 2-D linear array on the device: the
// per-pixel part of pycbf's make_tabs_and_apods_2D
// (pycbf/helpers/__full_aperture_2D__.py:3-203), so that Tabbed
// beamformers get their tables without a host round trip.
//
// The reference computes in float64 with NumPy; every per-pixel quantity
// here uses the same float64 operations in the same order (compiled with
// -ffp-contract=off), then rounds to float32 as the beamformer's table
// conversion does.  Per-transmit constants that need transcendental
// functions (steering sines, the synthetic point, the arccos acceptance
// threshold) come from the host, computed by the reference's own NumPy
// calls.  One lane per (element, pixel); rows [0, ntx) are transmits and
// rows [ntx, ntx+nrx) receive elements.
#include <math.h>

#include "common.h"

namespace pycbf {

__global__ __launch_bounds__(256) void k_tabs_2d(int mode, int ntx, const double *__restrict__ txc,
                                                 int nrx, const double *__restrict__ xele,
                                                 const double *__restrict__ xout, int64_t nx,
                                                 const double *__restrict__ zout, int64_t nz,
                                                 double cm, double two_fnum,
                                                 float *__restrict__ tautx,
                                                 float *__restrict__ apodtx,
                                                 float *__restrict__ taurx,
                                                 float *__restrict__ apodrx)
{
    const int64_t np = nx * nz;
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const int e = blockIdx.y;
    // pixel p = ix * nz + iz (meshgrid(xout, zout, indexing='ij') flattened)
    const double x = xout[p / nz], z = zout[p % nz];
    double tau, apod;
    if (e >= ntx) {
        // receive: |P - E| / cm, f-number cone |dx| - dz / (2F) < 0 (:24-33)
        const int r = e - ntx;
        const double dx = x - xele[r], dz = z - 0.0;
        tau = sqrt(dx * dx + dz * dz) / cm;
        apod = (fabs(dx) - dz / two_fnum < 0.0) ? 1.0 : 0.0;
        const int64_t o = (int64_t)r * np + p;
        if (taurx) taurx[o] = (float)tau;
        if (apodrx) apodrx[o] = (float)apod;
        return;
    }
    const double *c = txc + (int64_t)e * 8;
    if (mode == PYCBF_TABS_PLANE) {
        // (P - O) . N / cm, inside the aperture's projection (:150-166);
        // c = {ox, oz, nx, nz, npx, npz, x_first_element, x_last_element}
        const double dx = x - c[0], dz = z - c[1];
        tau = (dx * c[2] + dz * c[3]) / cm;
        const double neg = (x - c[6]) * c[4] + (z - 0.0) * c[5];
        const double pos = (x - c[7]) * c[4] + (z - 0.0) * c[5];
        apod = (neg <= 0.0 && pos >= 0.0) ? 1.0 : 0.0;
    } else if (mode == PYCBF_TABS_DIVERGENT) {
        // |P - O| / cm - tau0, no mask (:71-82); c = {ox, oz, tau0}
        const double dx = x - c[0], dz = z - c[1];
        tau = sqrt(dx * dx + dz * dz) / cm - c[2];
        apod = 1.0;
    } else {
        // sign(proj) |P - O| / cm + t0 and arccos(|proj| / max(|P - O|, 1e-12))
        // < ala (:116-125); c = {ox, oz, nx, nz, t0, vstar}: arccos(v) < ala
        // holds exactly for vstar <= v <= 1 (NumPy's arccos is monotone; vstar
        // is found on the host by bisection over doubles)
        const double dx = x - c[0], dz = z - c[1];
        const double proj = dx * c[2] + dz * c[3];
        const double mag = sqrt(dx * dx + dz * dz);
        const double sgn = proj > 0.0 ? 1.0 : (proj < 0.0 ? -1.0 : (proj == 0.0 ? 0.0 : proj));
        tau = sgn * mag / cm + c[4];
        const double m = mag < 1e-12 ? 1e-12 : mag;
        const double v = fabs(proj) / m;
        apod = (v >= c[5] && v <= 1.0) ? 1.0 : 0.0;
    }
    const int64_t o = (int64_t)e * np + p;
    if (tautx) tautx[o] = (float)tau;
    if (apodtx) apodtx[o] = (float)apod;
}

int launch_tabs_2d(int mode, int ntx, const double *txc, int nrx, const double *xele,
                   const double *xout, int64_t nx, const double *zout, int64_t nz, double cm,
                   double fnum, float *tautx, float *apodtx, float *taurx, float *apodrx,
                   hipStream_t s)
{
    const int64_t np = nx * nz;
    if (np <= 0 || ntx + nrx <= 0) return 0;
    dim3 grid((unsigned)((np + 255) / 256), (unsigned)(ntx + nrx));
    hipLaunchKernelGGL(k_tabs_2d, grid, dim3(256), 0, s, mode, ntx, txc, nrx, xele, xout, nx,
                       zout, nz, cm, 2 * fnum, tautx, apodtx, taurx, apodrx);
    return check_launch("tabs 2d");
}

}  // namespace pycbf
