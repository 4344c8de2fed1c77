// sph_nnps.hip -- cell-list neighbour search rebuilt as an on-GPU radix sort.
//
// Replaces PySPH's NNPS + DomainManager(periodic_mode="minimum_image") as
// used by the reference (application.py:63-78, scheme.py:71 update_nnps).
// Definitions (shared with the CPU oracle, which restates them):
//   c      = R * h_max of the instance, doubled while prod(n_a) > max(1, 2 N)
//   n_a    = max(1, floor(L_a / c)),  c_eff_a = L_a / n_a
//   key    = cell_base + ix + n_x (iy + n_y iz),  i_a = clamp(floor((x_a-lo_a)/c_eff_a))
//   sort   = stable by key (storage index breaks ties)
//   pairs  = r^2 < (R max(h_i, h_j))^2 with the minimum image
// All of it is HBM-bound integer work: coalesced key writes, one LDS
// histogram per wave tile and a stable multisplit scatter per 8-bit digit.
#include <algorithm>

#include "sph_internal.h"

namespace sph {

// ---------------------------------------------------------------------------
// per-instance bookkeeping
// ---------------------------------------------------------------------------

__global__ void k_eval_begin(Inst* inst, int ninst, Ctl* ctl, int reset_dt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        ctl->all_done = 0;
        ctl->status = 0;
        ctl->max_nb = 0;
        ctl->iters_max = 0;
        ctl->pairs = 0;
        ctl->count = 0;
        ctl->act_cnt[0] = 0;
        ctl->act_cnt[1] = 0;
        ctl->overflow = 0;
    }
    if (k >= ninst) return;
    Inst& s = inst[k];
    s.hmax_bits = 0ull;
    s.hmax_next_bits = 0ull;
    s.hmax_conv_bits = 0ull;
    if (reset_dt) s.dtp_bits = 0x7FF0000000000000ull;  // +inf
    s.iters = 0;
    s.unconv = 0;
    s.done = 0;
    s.status = 0;
}

// max h per instance.  Same-address atomics serialise, so the grid is small
// (kHmaxBlocks blocks, each over one contiguous chunk): a thread keeps a
// running max for the instance it is in and flushes it when the instance
// changes (instances are contiguous), then waves and the block combine what
// is left into one atomic per block for a single instance.
constexpr int kHmaxBlocks = 128;

__global__ void k_hmax(const double* __restrict__ h, const int32_t* __restrict__ inst_of, Inst* inst,
                       int64_t n, int next, int64_t chunk) {
    __shared__ unsigned long long wmax[kBlock / kWave];
    __shared__ int wk[kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = min(c0 + chunk, n);
    int kcur = -1;
    unsigned long long b = 0ull;
    for (int64_t i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
        const int k = inst_of[i];
        if (k != kcur) {
            if (kcur >= 0) atomicMax(next ? &inst[kcur].hmax_next_bits : &inst[kcur].hmax_bits, b);
            kcur = k;
            b = 0ull;
        }
        const unsigned long long v = dbits(h[i]);
        b = v > b ? v : b;
    }
    // threads with no element carry kcur = -1 and b = 0
    int k0 = kcur;
    for (int off = 32; off > 0; off >>= 1) k0 = max(k0, __shfl_xor(k0, off));
    const bool uniform = __all(kcur == k0 || kcur < 0);
    if (uniform) {
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(b, off);
            b = o > b ? o : b;
        }
    } else if (kcur >= 0) {
        atomicMax(next ? &inst[kcur].hmax_next_bits : &inst[kcur].hmax_bits, b);
    }
    if (lane == 0) {
        wmax[wave] = b;
        wk[wave] = uniform ? k0 : -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < kBlock / kWave;) {
            const int k = wk[q];
            unsigned long long m = wmax[q];
            int r = q + 1;
            for (; r < kBlock / kWave && wk[r] == k; ++r) m = wmax[r] > m ? wmax[r] : m;
            if (k >= 0) atomicMax(next ? &inst[k].hmax_next_bits : &inst[k].hmax_bits, m);
            q = r;
        }
    }
}

// Grid set-up: one thread per instance, then a serial prefix of cell bases.
__global__ void k_grid(Inst* inst, int ninst, Ctl* ctl) {
    int k = threadIdx.x;
    for (; k < ninst; k += blockDim.x) {
        Inst& s = inst[k];
        int64_t nk = s.off1 - s.off0;
        int64_t cap = 2 * nk > 1 ? 2 * nk : 1;
        double R = kernel_radius(s.kernel);
        double c = R * bitsd(s.hmax_bits);
        double lmax = 0.0;
        for (int a = 0; a < s.dim; ++a) lmax = fmax(lmax, s.L[a]);
        if (!(c > 0.0) || !(c <= lmax)) c = lmax;  // h_max of 0, NaN or above the box: one cell
        for (int guard = 0; guard < 80; ++guard) {
            int64_t prod = 1;
            for (int a = 0; a < 3; ++a) {
                if (a < s.dim) {
                    double nf = floor(s.L[a] / c);
                    int na = nf < 1.0 ? 1 : (nf > 1.0e9 ? 1000000000 : (int)nf);
                    s.n[a] = na;
                    prod *= na;
                    if (prod > cap) break;
                } else {
                    s.n[a] = 1;
                }
            }
            if (prod <= cap) break;
            c = 2.0 * c;
        }
        for (int a = 0; a < 3; ++a) s.ceff[a] = a < s.dim ? s.L[a] / s.n[a] : 1.0;
        s.ncell = (int64_t)s.n[0] * s.n[1] * s.n[2];
        stencil_from_hmax(s);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t base = 0;
        for (int q = 0; q < ninst; ++q) {
            inst[q].cell_base = base;
            base += inst[q].ncell;
        }
        ctl->total_cells = base;
    }
}

__global__ void k_keys(Fields F, const int32_t* __restrict__ inst_of, const Inst* __restrict__ inst,
                       uint32_t* __restrict__ keys, int32_t* __restrict__ vals, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Inst& s = inst[inst_of[i]];
    int ix = cell_coord(F.f[SPHMHD_F_X][i], s.lo[0], s.ceff[0], s.n[0]);
    int iy = cell_coord(F.f[SPHMHD_F_Y][i], s.lo[1], s.ceff[1], s.n[1]);
    int iz = s.dim == 3 ? cell_coord(F.f[SPHMHD_F_Z][i], s.lo[2], s.ceff[2], s.n[2]) : 0;
    int64_t local = ix + (int64_t)s.n[0] * (iy + (int64_t)s.n[1] * iz);
    keys[i] = (uint32_t)(s.cell_base + local);
    vals[i] = (int32_t)i;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort, 8-bit digits, wave-sized tiles
// ---------------------------------------------------------------------------

constexpr int RS_TILE = kRadixTile;          // keys per wave tile
constexpr int RS_ITEMS = RS_TILE / kWave;
constexpr int RS_WAVES = kBlock / kWave;

// A pass whose digit lies above the highest bit of the largest key (the cell
// count is only known on the device) is an identity: hist/scan return at
// once and the scatter becomes a copy, so the result always ends in the same
// buffer whatever the number of live passes.
__device__ __forceinline__ bool rs_pass_dead(const Ctl* ctl, int shift) {
    const int64_t top = ctl->total_cells - 1;
    return shift > 0 && top >= 0 && (top >> shift) == 0;
}

__global__ void k_rs_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                          uint32_t* __restrict__ hist, int64_t tiles, const Ctl* ctl) {
    if (rs_pass_dead(ctl, shift)) return;
    __shared__ uint32_t h[RS_WAVES][256];
    int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    for (int d = lane; d < 256; d += kWave) h[wave][d] = 0;
    __syncthreads();
    int64_t tile = (int64_t)blockIdx.x * RS_WAVES + wave;
    if (tile < tiles) {
        int64_t base = tile * RS_TILE;
        for (int it = 0; it < RS_ITEMS; ++it) {
            int64_t idx = base + it * kWave + lane;
            if (idx < n) atomicAdd(&h[wave][(keys[idx] >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    if (tile < tiles)
        for (int d = lane; d < 256; d += kWave) hist[(int64_t)d * tiles + tile] = h[wave][d];
}

// exclusive scan of one digit's row of tile counts; row totals to tot[d]
__global__ void k_rs_scan_rows(uint32_t* __restrict__ hist, int64_t tiles, uint32_t* __restrict__ tot,
                               int shift, const Ctl* ctl) {
    if (rs_pass_dead(ctl, shift)) return;
    __shared__ uint32_t part[kBlock];
    uint32_t* row = hist + (int64_t)blockIdx.x * tiles;
    int64_t per = (tiles + kBlock - 1) / kBlock;
    int64_t a = threadIdx.x * per, b = a + per < tiles ? a + per : tiles;
    uint32_t s = 0;
    for (int64_t q = a; q < b; ++q) s += row[q];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (int64_t q = a; q < b; ++q) {
        uint32_t c = row[q];
        row[q] = run;
        run += c;
    }
    if (threadIdx.x == kBlock - 1) tot[blockIdx.x] = part[kBlock - 1];
}

__global__ void k_rs_scan_digits(uint32_t* __restrict__ tot, int shift, const Ctl* ctl) {
    if (rs_pass_dead(ctl, shift)) return;
    __shared__ uint32_t part[256];
    uint32_t v = tot[threadIdx.x];
    part[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t o = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += o;
        __syncthreads();
    }
    tot[threadIdx.x] = part[threadIdx.x] - v;
}

__global__ void k_rs_scatter(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                             uint32_t* __restrict__ kout, int32_t* __restrict__ vout, int64_t n, int shift,
                             const uint32_t* __restrict__ hist, const uint32_t* __restrict__ dbase, int64_t tiles,
                             const Ctl* ctl) {
    __shared__ uint32_t run[RS_WAVES][256];
    int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    int64_t tile = (int64_t)blockIdx.x * RS_WAVES + wave;
    if (rs_pass_dead(ctl, shift)) {
        if (tile < tiles)
            for (int it = 0; it < RS_ITEMS; ++it) {
                int64_t idx = tile * RS_TILE + it * kWave + lane;
                if (idx < n) {
                    kout[idx] = kin[idx];
                    vout[idx] = vin[idx];
                }
            }
        return;
    }
    for (int d = lane; d < 256; d += kWave)
        run[wave][d] = tile < tiles ? dbase[d] + hist[(int64_t)d * tiles + tile] : 0u;
    __syncthreads();
    if (tile >= tiles) return;
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int64_t base = tile * RS_TILE;
    for (int it = 0; it < RS_ITEMS; ++it) {
        int64_t idx = base + it * kWave + lane;
        bool valid = idx < n;
        uint32_t key = valid ? kin[idx] : 0u;
        int32_t val = valid ? vin[idx] : 0;
        uint32_t dg = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (dg >> b) & 1u;
            unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t before = (uint32_t)__popcll(peers & lt);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t pos = run[wave][dg] + before;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            kout[pos] = key;
            vout[pos] = val;
            if (before == 0) run[wave][dg] = pos + cnt;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// cell table and packs
// ---------------------------------------------------------------------------

// cell table: cstart[c] = first sorted position with key >= c and cend[c] =
// first with key > c, for every cell (binary search of the sorted keys), so
// empty cells still bound their neighbours' ranges
__device__ __forceinline__ int32_t lower_bound_key(const uint32_t* __restrict__ skey, int64_t n, uint32_t k) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (skey[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return (int32_t)lo;
}

__global__ void k_cell_bounds(const uint32_t* __restrict__ skey, int32_t* __restrict__ cstart,
                              int32_t* __restrict__ cend, int64_t n, const Ctl* ctl, int64_t cap) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cap || c >= ctl->total_cells) return;
    cstart[c] = lower_bound_key(skey, n, (uint32_t)c);
    cend[c] = lower_bound_key(skey, n, (uint32_t)(c + 1));
}

__global__ void k_pack_density(Fields F, const int32_t* __restrict__ perm, double4* __restrict__ dpack, int64_t n,
                               int dim) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int32_t i = perm[t];
    double4 r;
    r.x = F.f[SPHMHD_F_X][i];
    r.y = F.f[SPHMHD_F_Y][i];
    r.z = dim == 3 ? F.f[SPHMHD_F_Z][i] : 0.0;
    r.w = F.f[SPHMHD_F_M][i];
    dpack[t] = r;
}

// positions and h only, for neighbour exports outside an evaluation
__global__ void k_pack_xh(Fields F, const int32_t* __restrict__ perm, double4* __restrict__ xh, int64_t n, int dim) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int32_t i = perm[t];
    double4 r;
    r.x = F.f[SPHMHD_F_X][i];
    r.y = F.f[SPHMHD_F_Y][i];
    r.z = dim == 3 ? F.f[SPHMHD_F_Z][i] : 0.0;
    r.w = F.f[SPHMHD_F_H][i];
    xh[t] = r;
}

// before re-running a force pass: forget the partial pass's dt minimum and
// warning bits (density warnings are kept)
__global__ void k_reset_force_outputs(Inst* inst, int ninst, Ctl* ctl) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) ctl->status = 0;
    if (k < ninst) {
        inst[k].dtp_bits = 0x7FF0000000000000ull;
        inst[k].status &= ~2;
    }
}

__global__ void k_reset_nb_stats(Ctl* ctl) {
    ctl->max_nb = 0;
    ctl->pairs = 0;
    ctl->overflow = 0;
}

// ---------------------------------------------------------------------------
// neighbour lists for the force pass, column-major (nb[k * n + t]): the force
// kernel's wave then reads one contiguous 256-byte row per neighbour slot
// (measured faster than per-thread rows, whose lines do not stay in L1)
// ---------------------------------------------------------------------------

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_nb_build(const double4* __restrict__ xh, const uint32_t* __restrict__ skey,
                                                     const int32_t* __restrict__ cstart, const int32_t* __restrict__ cend,
                                                     const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of,
                                                     int32_t* __restrict__ nb, int32_t* __restrict__ nbcnt, int64_t n,
                                                     int stride, Ctl* ctl) {
    // 2-D blocks stage their candidates in LDS (measured 7-20 % faster); in
    // 3-D the staged runs rarely fit and the LDS would cost occupancy
    constexpr int kCap = DIM == 2 ? kStageCap : 1;
    __shared__ StageHdr H;
    __shared__ double4 stage[kCap];
    const int64_t t0 = xcd_tile((n + kBlock - 1) / kBlock) * kBlock;
    if (t0 >= n) return;  // whole block past the end (the block barriers below need every wave)
    if constexpr (DIM == 2) {
        if (threadIdx.x < kWave) stage_plan<DIM, kWave>(H, skey, cstart, cend, inst, inst_of, t0, kBlock, n, kStageCap);
    } else {
        if (threadIdx.x == 0) H.on = 0;
    }
    __syncthreads();
    stage_fill(H, stage, xh);
    __syncthreads();

    const int64_t t = t0 + threadIdx.x;
    const bool valid = t < n;
    int cnt = 0;
    if (valid) {
        const Inst& s = inst[inst_of[t]];
        const double R = kernel_radius(s.kernel);
        const double4 me = xh[t];
        const double xi = me.x, yi = me.y, zi = me.z, hi = me.w;
        const double L0 = s.L[0], L1 = s.L[1], L2 = s.L[2];
        const double H0 = s.halfL[0], H1 = s.halfL[1], H2 = s.halfL[2];
        const int per = s.periodic;
        for_each_candidate<DIM>(s, (int64_t)skey[t], cstart, cend, H, stage, xh, [&](int j, const double4& o) {
            const double dx = min_image(xi - o.x, L0, H0, per);
            const double dy = min_image(yi - o.y, L1, H1, per);
            const double dz = DIM == 3 ? min_image(zi - o.z, L2, H2, per) : 0.0;
            const double r2 = r2_exact(dx, dy, dz);
            const double rad = R * fmax(hi, o.w);
            if (r2 < rad * rad) {
                if (cnt < stride) nb[(int64_t)cnt * n + t] = j;
                ++cnt;
            }
        });
        nbcnt[t] = cnt;
    }
    // longest list and pair count: wave, then block, then one atomic each
    // (same-address atomics serialise, so one per block rather than per wave)
    int m = cnt;
    long long pc = cnt;
    for (int off = 32; off > 0; off >>= 1) {
        m = max(m, __shfl_xor(m, off));
        pc += __shfl_xor(pc, off);
    }
    __shared__ int bm[kBlock / kWave];
    __shared__ long long bpc[kBlock / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) {
        bm[threadIdx.x / kWave] = m;
        bpc[threadIdx.x / kWave] = pc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / kWave; ++w) {
            m = max(m, bm[w]);
            pc += bpc[w];
        }
        if (pc) {
            atomicMax(&ctl->max_nb, m);
            atomicAdd((unsigned long long*)&ctl->pairs, (unsigned long long)pc);
            if (m > stride) atomicOr(&ctl->overflow, 1);
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void launch_eval_begin(Dev& d, bool reset_dt) {
    hipLaunchKernelGGL(k_eval_begin, dim3(nblk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, d.ninst, d.ctl,
                       reset_dt ? 1 : 0);
}

void launch_pack_xh(Dev& d) {
    if (d.n == 0) return;
    hipLaunchKernelGGL(k_pack_xh, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.perm, (double4*)d.dpack,
                       d.n, d.dim);
}

void launch_reset_force_outputs(Dev& d) {
    hipLaunchKernelGGL(k_reset_force_outputs, dim3(nblk(d.ninst)), dim3(kBlock), 0, d.stream, d.inst, d.ninst, d.ctl);
}

void launch_reset_nb_stats(Dev& d) {
    hipLaunchKernelGGL(k_reset_nb_stats, dim3(1), dim3(1), 0, d.stream, d.ctl);
}

void launch_hmax(Dev& d, bool next) {
    if (d.n == 0) return;
    const int64_t blocks = std::min<int64_t>(kHmaxBlocks, nblk(d.n));
    const int64_t chunk = (d.n + blocks - 1) / blocks;
    hipLaunchKernelGGL(k_hmax, dim3((unsigned)blocks), dim3(kBlock), 0, d.stream, d.fields.f[SPHMHD_F_H], d.inst_of,
                       d.inst, d.n, next ? 1 : 0, chunk);
}

void launch_grid(Dev& d) {
    hipLaunchKernelGGL(k_grid, dim3(1), dim3(kBlock), 0, d.stream, d.inst, d.ninst, d.ctl);
}

void launch_keys_sort_ranges(Dev& d) {
    if (d.n == 0) return;
    // the keys start in whichever buffer makes the last pass land in d.keys,
    // so the buffers (and a captured graph's arguments) never swap
    const int passes = (d.key_bits + 7) / 8;
    uint32_t *ka = (passes & 1) ? d.keys_alt : d.keys, *kb = (passes & 1) ? d.keys : d.keys_alt;
    int32_t *va = (passes & 1) ? d.perm_alt : d.perm, *vb = (passes & 1) ? d.perm : d.perm_alt;
    hipLaunchKernelGGL(k_keys, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst_of, d.inst, ka, va, d.n);
    int64_t tiles = (d.n + RS_TILE - 1) / RS_TILE;
    unsigned blocks = (unsigned)((tiles + RS_WAVES - 1) / RS_WAVES);
    for (int shift = 0; shift < d.key_bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(blocks), dim3(kBlock), 0, d.stream, ka, d.n, shift, d.rs_hist, tiles,
                           d.ctl);
        hipLaunchKernelGGL(k_rs_scan_rows, dim3(256), dim3(kBlock), 0, d.stream, d.rs_hist, tiles, d.rs_base, shift,
                           d.ctl);
        hipLaunchKernelGGL(k_rs_scan_digits, dim3(1), dim3(256), 0, d.stream, d.rs_base, shift, d.ctl);
        hipLaunchKernelGGL(k_rs_scatter, dim3(blocks), dim3(kBlock), 0, d.stream, ka, va, kb, vb, d.n, shift,
                           d.rs_hist, d.rs_base, tiles, d.ctl);
        uint32_t* tk = ka; ka = kb; kb = tk;
        int32_t* tv = va; va = vb; vb = tv;
    }
    hipLaunchKernelGGL(k_cell_bounds, dim3(nblk(d.cell_cap)), dim3(kBlock), 0, d.stream, d.keys, d.cstart, d.cend, d.n,
                       d.ctl, d.cell_cap);
}

void launch_pack_density(Dev& d) {
    if (d.n == 0) return;
    hipLaunchKernelGGL(k_pack_density, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.perm,
                       (double4*)d.dpack, d.n, d.dim);
}

void launch_nb_build(Dev& d) {
    if (d.n == 0) return;
    if (d.dim == 2)
        hipLaunchKernelGGL(k_nb_build<2>, dim3(xcd_grid(nblk(d.n))), dim3(kBlock), 0, d.stream, (const double4*)d.dpack, d.keys,
                           d.cstart,
                           d.cend, d.inst, d.inst_of, d.nb, d.nbcnt, d.n, d.nb_stride, d.ctl);
    else
        hipLaunchKernelGGL(k_nb_build<3>, dim3(xcd_grid(nblk(d.n))), dim3(kBlock), 0, d.stream, (const double4*)d.dpack, d.keys,
                           d.cstart,
                           d.cend, d.inst, d.inst_of, d.nb, d.nbcnt, d.n, d.nb_stride, d.ctl);
}

}  // namespace sph
