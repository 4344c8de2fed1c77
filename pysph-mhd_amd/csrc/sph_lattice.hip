// sph_lattice.hip -- on-GPU initial positions: the close-packed (HCP,
// ABC-stacked) lattice of the reference's cases (src/cases/lattice.py:4-34).
//
// One thread per site, in (i_x, i_y, i_z) C order.  The arithmetic and its
// order are the reference's (numpy evaluates left to right with no fused
// multiply-add); contraction is switched off in the kernel (HIP's __dmul_rn /
// __dadd_rn are plain operators that the compiler would still fuse into
// v_fmac_f64), so the positions equal the reference's bit for bit.
#include <cmath>
#include <string>

#include "sph_internal.h"

namespace sph {

__global__ void k_hcp_lattice(int64_t nx, int64_t ny, int64_t nz, double xbase, double ymin, double zmin, double dx,
                              double dy, double dz, double* __restrict__ x, double* __restrict__ y,
                              double* __restrict__ z) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = ny * nz;
    if (i >= nx * per) return;
    const int64_t k = i / per, rem = i - k * per;
    const int64_t ly = rem / nz + 1, lz = rem - (ly - 1) * nz + 1;
    const int ypar = (int)(ly % 2), zmod = (int)(lz % 3);
    const bool shifted = (zmod == 0 && ypar == 0) || (zmod == 2 && ypar == 1) || (zmod == 1 && ypar == 0);
    const double xoff = shifted ? 0.5 : 0.0;
    const double yoff = zmod == 0 ? 2.0 / 3.0 : (zmod == 2 ? 1.0 / 3.0 : 0.0);
    x[i] = xbase + ((double)k + xoff) * dx;
    y[i] = ymin + ((((double)ly - 1.0) + 1.0 / 6.0) + yoff) * dy;
    z[i] = zmin + ((double)lz - 0.5) * dz;
}

}  // namespace sph

using namespace sph;

extern "C" int sphmhd_hcp_lattice(const double* bounds, double spacing, double* x, double* y, double* z,
                                  int64_t cap, int64_t* count, void* hip_stream) {
    if (!bounds || !count) return report_error(SPHMHD_EINVAL, "null argument");
    if (!(spacing > 0.0)) return report_error(SPHMHD_EINVAL, "spacing must be positive");
    const double xmin = bounds[0], xmax = bounds[1], ymin = bounds[2], ymax = bounds[3];
    const double zmin = bounds[4], zmax = bounds[5];
    if (!(xmax > xmin) || !(ymax > ymin) || !(zmax > zmin)) return report_error(SPHMHD_EINVAL, "empty bounds");
    // lattice.py:6-11: each quotient stepped one ulp down before the ceiling
    const double qx = std::ceil(std::nextafter((xmax - xmin) / spacing, -INFINITY));
    const double qy = std::ceil(std::nextafter((ymax - ymin) / (spacing * std::sqrt(3.0 / 4.0)), -INFINITY));
    const double qz = std::ceil(std::nextafter((zmax - zmin) / (spacing * std::sqrt(6.0) / 3.0), -INFINITY));
    if (!(qx < 1.0e7) || !(qy < 1.0e7) || !(qz < 1.0e7)) return report_error(SPHMHD_EINVAL, "lattice too large");
    const int64_t nx = (int64_t)qx, ny = 2 * ((int64_t)qy / 2), nz = 3 * ((int64_t)qz / 3);
    const int64_t n = nx > 0 && ny > 0 && nz > 0 ? nx * ny * nz : 0;
    *count = n;
    if (n > cap) return report_error(SPHMHD_ECAP, "lattice has " + std::to_string(n) + " sites, capacity " +
                                                      std::to_string(cap));
    if (n == 0) return 0;
    if (!x || !y || !z) return report_error(SPHMHD_EINVAL, "null output");
    const double dx = (xmax - xmin) / (double)nx, dy = (ymax - ymin) / (double)ny, dz = (zmax - zmin) / (double)nz;
    const double xbase = xmin + 0.25 * spacing;
    const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_hcp_lattice, dim3(blocks), dim3(kBlock), 0, (hipStream_t)hip_stream, nx, ny, nz, xbase,
                       ymin, zmin, dx, dy, dz, x, y, z);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return report_error(SPHMHD_EHIP, hipGetErrorString(e));
    return 0;
}
