// sph_cases.hip -- initial fields of the benchmark workloads and of the
// reference's cases, computed on the GPU (reference: src/cases/*.py).
//
// One thread per particle evaluates the case's analytic fields from its
// position: the benchmark lattices (square and cubic, particles at cell
// centres) generate their positions too; the reference's cases take
// positions the caller placed (the close-packed lattice of sphmhd_hcp_lattice,
// masked and joined where a case combines two lattices or several jets).
// Every constant that does not depend on the particle (masses, smoothing
// lengths, densities, trigonometry of fixed angles) comes from the host in
// the argument block, computed there exactly as the host builders do, and
// the per-particle arithmetic follows the host builders' expressions in
// their order with contraction off, so the fields agree with the host
// builders' to the last bit except through the transcendental functions (sin,
// cos, exp, tanh, cosh, pow), whose GPU and host library versions may differ
// in the last place.
#include <cmath>
#include <string>

#include "sph_internal.h"

namespace sph {

struct CaseArgs {
    int32_t id;
    int32_t pad;
    int64_t n;       // particles written
    int64_t n_aux;   // particles per jet (jets)
    double p[48];
};

// a row of the bound layout (rows other than x..Bz, alpha1 may be null)
__device__ __forceinline__ void put(const Fields& F, int f, int64_t i, double v) {
    if (F.f[f]) F.f[f][i] = v;
}

__device__ __forceinline__ void put_fluid(const Fields& F, int64_t i, double u, double v, double w, double rho,
                                          double h, double m, double e, double bx, double by, double bz,
                                          double alpha1) {
    put(F, SPHMHD_F_U, i, u);
    put(F, SPHMHD_F_V, i, v);
    put(F, SPHMHD_F_W, i, w);
    put(F, SPHMHD_F_RHO, i, rho);
    put(F, SPHMHD_F_H, i, h);
    put(F, SPHMHD_F_M, i, m);
    put(F, SPHMHD_F_E, i, e);
    put(F, SPHMHD_F_BX, i, bx);
    put(F, SPHMHD_F_BY, i, by);
    put(F, SPHMHD_F_BZ, i, bz);
    put(F, SPHMHD_F_ALPHA1, i, alpha1);
}

__global__ void k_case_fields(Fields F, CaseArgs a, const double* __restrict__ aux) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double* p = a.p;
    const double pi = 3.14159265358979323846;
    const double two_pi = 2.0 * pi;
    double* const* f = F.f;
    switch (a.id) {
    case SPHMHD_CASE_ORSZAG_TANG_2D: {
        // p: nx, dx, rho0, m, h, e, b0, alpha0
        const int64_t nx = (int64_t)p[0];
        const double dx = p[1];
        const double x = ((double)(i % nx) + 0.5) * dx, y = ((double)(i / nx) + 0.5) * dx;
        put(F, SPHMHD_F_X, i, x);
        put(F, SPHMHD_F_Y, i, y);
        put(F, SPHMHD_F_Z, i, 0.0);
        const double b0 = p[6];
        put_fluid(F, i, -sin(two_pi * y), sin(two_pi * x), 0.0, p[2], p[4], p[3], p[5], -b0 * sin(two_pi * y),
                  b0 * sin(2.0 * two_pi * x), 0.0, p[7]);
        break;
    }
    case SPHMHD_CASE_ROTOR_2D: {
        // p: nx, dx, h, gamma, b
        const int64_t nx = (int64_t)p[0];
        const double dx = p[1];
        const double x = ((double)(i % nx) + 0.5) * dx, y = ((double)(i / nx) + 0.5) * dx;
        put(F, SPHMHD_F_X, i, x);
        put(F, SPHMHD_F_Y, i, y);
        put(F, SPHMHD_F_Z, i, 0.0);
        const double r0 = 0.1, r1 = 0.115, omega = 20.0;
        const double rx = x - 0.5, ry = y - 0.5;
        const double r = sqrt(rx * rx + ry * ry);
        const double ft = fmin(fmax((r1 - r) / (r1 - r0), 0.0), 1.0);
        const bool inside = r < r0;
        const double rho = inside ? 10.0 : 1.0 + 9.0 * ft;
        const double safe_r = r > 0.0 ? r : 1.0;
        const double vfac = inside ? omega : ft * omega * r0 / safe_r;
        put_fluid(F, i, -vfac * ry, vfac * rx, 0.0, rho, p[2], rho * dx * dx, 1.0 / ((p[3] - 1.0) * rho), p[4], 0.0,
                  0.0, 0.0);
        break;
    }
    case SPHMHD_CASE_BLAST_3D: {
        // p: nx, dx, m, h, gamma, b, r_blast^2
        const int64_t nx = (int64_t)p[0];
        const double dx = p[1];
        const double x = ((double)(i % nx) + 0.5) * dx;
        const double y = ((double)((i / nx) % nx) + 0.5) * dx;
        const double z = ((double)(i / (nx * nx)) + 0.5) * dx;
        put(F, SPHMHD_F_X, i, x);
        put(F, SPHMHD_F_Y, i, y);
        put(F, SPHMHD_F_Z, i, z);
        const double ax = x - 0.5, ay = y - 0.5, az = z - 0.5;
        const double r2 = ax * ax + ay * ay + az * az;
        const double pres = r2 < p[6] ? 100.0 : 1.0;
        put_fluid(F, i, 0.0, 0.0, 0.0, 1.0, p[3], p[2], pres / (p[4] - 1.0), p[5], p[5], 0.0, 0.0);
        break;
    }
    case SPHMHD_CASE_ALFVEN_2D: {
        // p: nx, dx, dy, amplitude, ca, sa, m, h, e
        const int64_t nx = (int64_t)p[0];
        const double dx = p[1], dy = p[2];
        const double x = ((double)(i % nx) + 0.5) * dx, y = ((double)(i / nx) + 0.5) * dy;
        put(F, SPHMHD_F_X, i, x);
        put(F, SPHMHD_F_Y, i, y);
        put(F, SPHMHD_F_Z, i, 0.0);
        const double ca = p[4], sa = p[5];
        const double ph = two_pi * (x * ca + y * sa);
        const double bperp = p[3] * sin(ph), bz = p[3] * cos(ph);
        put_fluid(F, i, -bperp * sa, bperp * ca, bz, 1.0, p[7], p[6], p[8], 1.0 * ca - bperp * sa, 1.0 * sa + bperp * ca,
                  bz, 0.0);
        break;
    }
    case SPHMHD_CASE_ORSTANG: {
        // p: xmin, ymin, rho, m, h, e, b0
        const double x = f[SPHMHD_F_X][i], y = f[SPHMHD_F_Y][i];
        const double b0 = p[6];
        put_fluid(F, i, -sin(2.0 * pi * (y - p[1])), sin(2.0 * pi * (x - p[0])), 0.0, p[2], p[4], p[3], p[5],
                  -b0 * sin(2.0 * pi * (y - p[1])), b0 * sin(4.0 * pi * (x - p[0])), 0.0, 0.0);
        break;
    }
    case SPHMHD_CASE_MHDROTOR: {
        // p: r0, omega_in, m, h_in, h_out, gamma1, b
        const double x = f[SPHMHD_F_X][i], y = f[SPHMHD_F_Y][i];
        const bool inside = sqrt(x * x + y * y) <= p[0];
        const double rho = inside ? 10.0 : 1.0;
        const double om = inside ? p[1] : 0.0;
        put_fluid(F, i, -om * y, om * x, 0.0, rho, inside ? p[3] : p[4], p[2], 1.0 / (p[5] * rho), p[6], 0.0, 0.0,
                  0.0);
        break;
    }
    case SPHMHD_CASE_MHDBLAST: {
        // p: rho, m, h, gamma1, b, r_blast^2
        const double x = f[SPHMHD_F_X][i], y = f[SPHMHD_F_Y][i], z = f[SPHMHD_F_Z][i];
        const double pres = x * x + y * y + z * z < p[5] ? 100.0 : 1.0;
        put_fluid(F, i, 0.0, 0.0, 0.0, p[0], p[2], p[1], pres / (p[0] * p[3]), p[4], 0.0, p[4], 0.0);
        break;
    }
    case SPHMHD_CASE_ALFVEN: {
        // p: m, h; positions rounded through float32 as the reference does
        const double x = (double)(float)f[SPHMHD_F_X][i];
        put(F, SPHMHD_F_X, i, x);
        put(F, SPHMHD_F_Y, i, (double)(float)f[SPHMHD_F_Y][i]);
        put(F, SPHMHD_F_Z, i, (double)(float)f[SPHMHD_F_Z][i]);
        const double phase = 2.0 * pi * (x + 1.5);
        const double by = 0.1 * sin(phase), bz = 0.1 * cos(phase);
        put_fluid(F, i, 0.0, by, bz, 1.0, p[1], p[0], 0.15, 1.0, by, bz, 0.0);
        break;
    }
    case SPHMHD_CASE_JADVECT: {
        // p: rho, m, h, w
        const double x = f[SPHMHD_F_X][i], y = f[SPHMHD_F_Y][i];
        const double r = sqrt(x * x + y * y);
        const bool inside = r < 0.3;
        const double bx = inside ? -0.001 * y / r : 0.0, by = inside ? 0.001 * x / r : 0.0;
        put_fluid(F, i, 2.0, 1.0, p[3], p[0], p[2], p[1], 1.5, bx, by, 0.0, 0.0);
        break;
    }
    case SPHMHD_CASE_MHDSHOCK: {
        // p: m, h_left, h_right, gamma1, x_wall_left, x_wall_right
        const double x = f[SPHMHD_F_X][i];
        const bool left = x <= 0.0;
        const double rho = left ? 1.0 : 0.125, pres = left ? 1.0 : 0.1;
        put_fluid(F, i, 0.0, 0.0, 0.0, rho, left ? p[1] : p[2], p[0], pres / (p[3] * rho), 0.75, left ? 1.0 : -1.0,
                  0.0, 0.0);
        if (F.itype) F.itype[i] = (x < p[4] || x > p[5]) ? 3 : 1;
        break;
    }
    case SPHMHD_CASE_MHDSINE: {
        // p: xmin, rho, m, h
        const double x = f[SPHMHD_F_X][i];
        put_fluid(F, i, 0.0, 0.0, 0.0, p[1], p[3], p[2], 7.5, 0.0, sin(2.0 * pi * (x - p[0])), 0.0, 0.0);
        break;
    }
    case SPHMHD_CASE_MHDVORTEX: {
        // p: rho, m, h, gamma1, 8 pi^2
        const double x = f[SPHMHD_F_X][i], y = f[SPHMHD_F_Y][i];
        const double g = exp(0.5 * (1.0 - x * x - y * y)) / (2.0 * pi);
        const double s = 1.0 - x * x - y * y;
        const double pres = 1.0 + exp(s) * (s - 1.0) / p[4];
        put_fluid(F, i, 1.0 - y * g, 1.0 + x * g, 1.0, p[0], p[2], p[1], pres / (p[0] * p[3]), -y * g, x * g, 0.0,
                  0.0);
        break;
    }
    case SPHMHD_CASE_MHDWAVE: {
        // p: rho, m, h, bx
        const double x = f[SPHMHD_F_X][i];
        const double q = x / (3.0 * p[2]);
        put_fluid(F, i, 0.01 * exp(-(q * q)), 0.0, 0.0, p[0], p[2], p[1], 0.75, p[3], 0.0, 0.0, 0.0);
        break;
    }
    case SPHMHD_CASE_JETS: {
        // p: rho, m, h, e, radius, inner, outer, taper, axial_speed,
        //    azimuthal_speed, axial_field, toroidal_field, toroidal_peak,
        //    then (cos, sin) of each jet's angle; aux: the single jet's
        //    (axial, cy, cz) rows of n_aux particles
        const int64_t per = a.n_aux;
        const int64_t jet = i / per, q = i - jet * per;
        const double axial = aux[q], cy = aux[per + q], cz = aux[2 * per + q];
        const double c = p[13 + 2 * jet], s = p[14 + 2 * jet];
        const double R = p[4], inner = p[5], outer = p[6], taper = p[7];
        const double r = sqrt(cy * cy + cz * cz);
        const double inv_r = r > 0.0 ? 1.0 / r : 0.0;
        const double ratio = r / R;
        const double rem = 1.0 - ratio * ratio;
        const double length = outer - inner;
        const double local = 0.5 * (inner + outer) - axial;
        const double ap = (local + 0.5 * length) / taper, am = (local - 0.5 * length) / taper;
        const double envelope = 0.5 * (tanh(ap) - tanh(am));
        const double cap = cosh(ap), cam = cosh(am);
        const double denv = 0.5 * (1.0 / (cap * cap) - 1.0 / (cam * cam)) / taper;
        const double rem3 = pow(rem, 3.0), rem4 = pow(rem, 4.0);
        // field (jet.py:95-135)
        const double bax = p[10] * (rem3 * (1.0 - 5.0 * (ratio * ratio))) * envelope;
        const double brad = -0.5 * p[10] * r * rem4 * denv;
        const double btor = p[11] * ratio * rem3 / p[12] * envelope;
        const double by = (brad * cy + btor * cz) * inv_r;
        const double bz = (brad * cz - btor * cy) * inv_r;
        // velocity
        const double vphi = p[9] * ratio * rem3 / p[12] * envelope;
        const double vy = vphi * cz * inv_r, vz = -vphi * cy * inv_r;
        put(F, SPHMHD_F_X, i, axial * c - cz * s);
        put(F, SPHMHD_F_Y, i, cy);
        put(F, SPHMHD_F_Z, i, axial * s + cz * c);
        put_fluid(F, i, -p[8] * c - vz * s, vy, -p[8] * s + vz * c, p[0], p[2], p[1], p[3], -bax * c - bz * s, by,
                  -bax * s + bz * c, 0.0);
        break;
    }
    default:
        break;
    }
}

}  // namespace sph

using namespace sph;

extern "C" int sphmhd_case_fields(int32_t case_id, const double* params, int32_t nparams, double* const* fields,
                                  int32_t* itype, int64_t n, int64_t n_aux, const double* aux_dev,
                                  void* hip_stream) {
    if (!fields || (nparams > 0 && !params)) return report_error(SPHMHD_EINVAL, "null argument");
    if (nparams < 0 || nparams > 48) return report_error(SPHMHD_EINVAL, "at most 48 case parameters");
    if (n < 0) return report_error(SPHMHD_EINVAL, "negative particle count");
    if (case_id < SPHMHD_CASE_ORSZAG_TANG_2D || case_id > SPHMHD_CASE_JETS)
        return report_error(SPHMHD_EINVAL, "unknown case " + std::to_string(case_id));
    for (int f : {SPHMHD_F_X, SPHMHD_F_Y, SPHMHD_F_Z, SPHMHD_F_U, SPHMHD_F_V, SPHMHD_F_W, SPHMHD_F_RHO, SPHMHD_F_H,
                  SPHMHD_F_M, SPHMHD_F_E, SPHMHD_F_BX, SPHMHD_F_BY, SPHMHD_F_BZ, SPHMHD_F_ALPHA1})
        if (!fields[f]) return report_error(SPHMHD_EINVAL, "field row " + std::to_string(f) + " is null");
    if (case_id == SPHMHD_CASE_JETS) {
        if (!aux_dev || n_aux < 1) return report_error(SPHMHD_EINVAL, "jets need the single jet's positions");
        const int64_t jets = n / n_aux;
        if (jets * n_aux != n || 13 + 2 * jets > nparams)
            return report_error(SPHMHD_EINVAL, "jets: n must be a multiple of n_aux with an angle per jet");
    }
    CaseArgs a{};
    a.id = case_id;
    a.n = n;
    a.n_aux = n_aux;
    for (int k = 0; k < nparams; ++k) a.p[k] = params[k];
    Fields F{};
    for (int k = 0; k < SPHMHD_NFIELDS; ++k) F.f[k] = fields[k];
    F.itype = itype;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_case_fields, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)hip_stream, F, a,
                       aux_dev);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return report_error(SPHMHD_EHIP, hipGetErrorString(e));
    return 0;
}
