// sph_totals_static.hip -- conserved totals with no allocation: the partial
// sums live in a file-scope device array instead of memory the library
// allocates, so sphmhd_totals_noalloc can be called where hipMalloc cannot
// (inside HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The array `scratch` holds kTotSlots partial-sum records of kTotStride
// doubles (one 128-byte line each) and, after them, the arrival counters of
// the launch's last-workgroup tail.  In host code `scratch` names a host-side
// stand-in, never the device memory: its device address is taken once per
// device with hipGetSymbolAddress (sphmhd_create calls scratch_setup with the
// context's device current and keeps the address in the context), and only
// that address is passed to kernels.  Every call on one GPU
// uses that GPU's copy of the array: the partial records are written before
// they are read in every launch and the arrival counters return to zero at
// the end of each launch, so nothing carries over between calls, but two
// calls that overlap on one GPU (two streams, two host threads) race.
#include <algorithm>
#include <mutex>
#include <string>

#include "sph_internal.h"

namespace sph {

constexpr int kTotSlots = 16384;   // partial-sum records: instances x blocks per instance
constexpr int kTotStride = 16;     // doubles per record (SPHMHD_NTOTALS used)
constexpr int kTotArriveDoubles = (kArriveInts * 4 + 7) / 8;
constexpr int kScratchDoubles = kTotSlots * kTotStride + kTotArriveDoubles;  // 2.0 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

__global__ void k_fill_i32_static(int32_t* p, int32_t v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Resolves the current device's copy of `scratch` (once per device) and
// clears its arrival counters; returns 0 or a status code.
int scratch_setup(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return report_error(SPHMHD_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return report_error(SPHMHD_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return report_error(SPHMHD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        int32_t* arrive = (int32_t*)((double*)p + (size_t)kTotSlots * kTotStride);
        hipLaunchKernelGGL(k_fill_i32_static, dim3((kArriveInts + 255) / 256), dim3(256), 0, nullptr, arrive, 0,
                           kArriveInts);
        e = hipDeviceSynchronize();
        if (e != hipSuccess) return report_error(SPHMHD_EHIP, std::string("scratch set-up: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    if (out) *out = g_scratch_ptr[dev];
    return 0;
}

// Block (b, k) sums particles of instance k in a grid-stride loop over its
// range and writes its partial record; the last block to arrive adds each
// instance's records in block order (a fixed order: the result does not
// depend on which block ends last) and writes the totals.
__global__ void __launch_bounds__(kBlock) k_totals_static(Fields F, const Inst* __restrict__ inst, int ninst,
                                                          double* __restrict__ part, int32_t* arrive,
                                                          double* __restrict__ out) {
    const int k = blockIdx.y, nbx = gridDim.x;
    const Inst& s = inst[k];
    double acc[SPHMHD_NTOTALS];
#pragma unroll
    for (int q = 0; q < SPHMHD_NTOTALS; ++q) acc[q] = 0.0;
    for (int64_t i = s.off0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.off1;
         i += (int64_t)nbx * blockDim.x) {
        double* const* f = F.f;
        const double m = f[SPHMHD_F_M][i], rho = f[SPHMHD_F_RHO][i];
        const double u = f[SPHMHD_F_U][i], v = f[SPHMHD_F_V][i], w = f[SPHMHD_F_W][i];
        const double bx = f[SPHMHD_F_BX][i], by = f[SPHMHD_F_BY][i], bz = f[SPHMHD_F_BZ][i];
        const double v2 = u * u + v * v + w * w;
        const double b2 = bx * bx + by * by + bz * bz;
        acc[SPHMHD_T_MASS] += m;
        acc[SPHMHD_T_MOMX] += m * u;
        acc[SPHMHD_T_MOMY] += m * v;
        acc[SPHMHD_T_MOMZ] += m * w;
        acc[SPHMHD_T_EKIN] += 0.5 * m * v2;
        acc[SPHMHD_T_ETHERM] += m * f[SPHMHD_F_E][i];
        acc[SPHMHD_T_EMAG] += 0.5 * m * b2 / (s.mu0 * rho);
        if (b2 > 0.0) acc[SPHMHD_T_DIVB] += (m / rho) * f[SPHMHD_F_H][i] * fabs(f[SPHMHD_F_DIVBSYMM][i]) / sqrt(b2);
        acc[SPHMHD_T_VOLUME] += m / rho;
    }
    // block sum: waves by shuffles, then the waves' sums in wave order
    __shared__ double wsum[kBlock / kWave][SPHMHD_NTOTALS];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < SPHMHD_NTOTALS; ++q) {
        double v = acc[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) wsum[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < SPHMHD_NTOTALS) {
        double v = 0.0;
        for (int w = 0; w < (int)blockDim.x / kWave; ++w) v += wsum[w][threadIdx.x];
        st_dev(&part[((int64_t)k * nbx + blockIdx.x) * kTotStride + threadIdx.x], v);
    }
    if (last_to_arrive(arrive)) {
        for (int idx = threadIdx.x; idx < ninst * SPHMHD_NTOTALS; idx += blockDim.x) {
            const int kk = idx / SPHMHD_NTOTALS, q = idx - kk * SPHMHD_NTOTALS;
            double v = 0.0;
            for (int b = 0; b < nbx; ++b) v += ld_dev(&part[((int64_t)kk * nbx + b) * kTotStride + q]);
            out[idx] = v;
        }
    }
}

}  // namespace sph

using namespace sph;

extern "C" int sphmhd_totals_noalloc(sphmhd_ctx* ctx, double* out_dev) {
    Dev* d = nullptr;
    int64_t longest = 0;
    if (!ctx) return report_error(SPHMHD_EINVAL, "null context");
    if (!out_dev) return report_error(SPHMHD_EINVAL, "null output");
    int rc = ctx_view(ctx, &d, &longest);
    if (rc) return rc;
    if (d->ninst > kTotSlots)
        return report_error(SPHMHD_ECAP, "sphmhd_totals_noalloc: " + std::to_string(d->ninst) +
                                             " instances exceed its maximum of " + std::to_string(kTotSlots));
    double* base = d->tot_scratch;  // this context's device: resolved by sphmhd_create
    if (!base) return report_error(SPHMHD_ESTATE, "sphmhd_totals_noalloc: scratch not set up");
    // blocks per instance: enough to cover the longest instance (at most 64),
    // within the kTotSlots partial records
    int bx = (int)std::min<int64_t>((longest + kBlock - 1) / kBlock, 64);
    bx = std::max(1, std::min(bx, kTotSlots / d->ninst));
    int32_t* arrive = (int32_t*)(base + (size_t)kTotSlots * kTotStride);
    hipLaunchKernelGGL(k_totals_static, dim3(bx, d->ninst), dim3(kBlock), 0, d->stream, d->fields, d->inst, d->ninst,
                       base, arrive, out_dev);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return report_error(SPHMHD_EHIP, std::string("sphmhd_totals_noalloc: ") + hipGetErrorString(e));
    return 0;
}
