// sph_nnps.hip -- cell-list neighbour search rebuilt as an on-GPU sort of the
// particles' cell keys (a one-digit radix sort: count, scan, place).
//
// Replaces PySPH's NNPS + DomainManager(periodic_mode="minimum_image") as
// used by the reference (application.py:63-78, scheme.py:71 update_nnps).
// Definitions (shared with the CPU oracle, which restates them):
//   c      = R * h_max of the instance, doubled while prod(n_a) > max(1, 2 N)
//   n_a    = max(1, floor(L_a / c)),  c_eff_a = L_a / n_a
//   key    = cell_base + ix + n_x (iy + n_y iz),  i_a = clamp(floor((x_a-lo_a)/c_eff_a))
//   sort   = stable by key (storage index breaks ties)
//   pairs  = r^2 < (R max(h_i, h_j))^2 with the minimum image
// All of it is HBM-bound integer work: one key per particle, one counter
// per cell, a scan over the cells and one scattered placement per particle.
#include <algorithm>
#include <climits>
#include <type_traits>

#include "sph_internal.h"

namespace sph {

// ---------------------------------------------------------------------------
// per-instance bookkeeping
// ---------------------------------------------------------------------------

// a standalone evaluation's start (not inside a pipelined run of steps): it
// also clears the abort flag
__global__ void k_eval_begin(Inst* inst, int ninst, Ctl* ctl, int reset_dt) {
    if (threadIdx.x == 0) ctl->abort = 0;
    eval_begin_body(inst, ninst, ctl, reset_dt);
}

// max h per instance.  Same-address atomics serialise, so the grid is small
// (kHmaxBlocks blocks, each over one contiguous chunk): a thread keeps a
// running max for the instance it is in and flushes it when the instance
// changes (instances are contiguous), then waves and the block combine what
// is left into one atomic per block for a single instance.  Each thread
// loads kHmaxBatch elements before folding them in, so the loads of a small
// grid are in flight together rather than one latency at a time.
constexpr int kHmaxBlocks = 128;
constexpr int kHmaxBatch = 8;


// h_max, h_min and the particles' box per instance (the kRed maxima of
// red_values), then (the launch's last workgroup) the cell grid
__global__ void k_hmax(Fields F, const int32_t* __restrict__ inst_of, Inst* inst, int64_t n, int64_t chunk,
                       int ninst, Ctl* ctl, int32_t* arrive) {
    if (ctl->abort) return;  // the whole launch: the step after an unfinished evaluation
    const double* __restrict__ h = F.f[SPHMHD_F_H];
    const double* __restrict__ x = F.f[SPHMHD_F_X];
    const double* __restrict__ y = F.f[SPHMHD_F_Y];
    const double* __restrict__ z = F.f[SPHMHD_F_Z];
    const int64_t c0 = (int64_t)blockIdx.x * chunk, c1 = min(c0 + chunk, n);
    uint32_t* hist = ctl->hist;  // h levels for the 3-D cell size
    __shared__ uint32_t lh[kHistSlots];
    if (hist) {
        hist_clear(lh);
        __syncthreads();
    }
    int kcur = -1;
    unsigned long long v[kRed] = {};
    auto flush = [&](int k, const unsigned long long* m) {
        for (int q = 0; q < kRed; ++q) atomicMax(&inst[k].red[q], m[q]);
    };
    for (int64_t i0 = c0 + threadIdx.x; i0 < c1; i0 += (int64_t)kHmaxBatch * blockDim.x) {
        int kk[kHmaxBatch];
        double hh[kHmaxBatch], xx[kHmaxBatch], yy[kHmaxBatch], zz[kHmaxBatch];
#pragma unroll
        for (int u = 0; u < kHmaxBatch; ++u) {
            const int64_t i = i0 + (int64_t)u * blockDim.x;
            const bool ok = i < c1;
            kk[u] = ok ? inst_of[i] : -1;
            hh[u] = ok ? h[i] : 0.0;
            xx[u] = ok ? x[i] : 0.0;
            yy[u] = ok ? y[i] : 0.0;
            zz[u] = ok && z ? z[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kHmaxBatch; ++u) {
            const int k = kk[u];
            if (k < 0) continue;
            if (k != kcur) {
                if (kcur >= 0) flush(kcur, v);
                kcur = k;
                for (int q = 0; q < kRed; ++q) v[q] = 0ull;
            }
            unsigned long long w[kRed];
            red_values(w, hh[u], xx[u], yy[u], zz[u]);
            if (hist) hist_add(lh, hh[u]);
            for (int q = 0; q < kRed; ++q) v[q] = w[q] > v[q] ? w[q] : v[q];
        }
    }
    // threads with no element carry kcur = -1 and v = 0
    block_max_runs<kRed>(kcur, v, [&](int k, const unsigned long long* m, bool) { flush(k, m); });
    if (hist) {
        __syncthreads();
        hist_flush(hist, lh);
    }
    if (last_to_arrive(arrive)) grid_setup(inst, ninst, ctl);
}

__global__ void k_grid(Inst* inst, int ninst, Ctl* ctl) { grid_setup(inst, ninst, ctl); }

// ---------------------------------------------------------------------------
// cell sort: one radix digit spanning the whole cell key (a counting sort).
// count -> two-level exclusive scan (cstart/cend of every cell) -> each
// particle placed in its cell's bucket -> each bucket put in storage order,
// which makes the sort stable (ties by storage index, as the definition
// asks).  Buckets are small (a cell is R h_max wide), so each particle of a
// bucket of up to kSmallBucket finds its place by its rank; larger buckets
// go to a workgroup bitonic sort (in LDS up to kBigLds, in global memory
// beyond: a one-cell grid holds every particle).  ccount returns to zero in
// the placement, ready for the next sort.
// ---------------------------------------------------------------------------

// Particles are visited in the previous sort's order (perm holds it; the
// identity before the first sort), where the particles of one cell are
// mostly adjacent: each run of equal keys in a wave does one atomic.
__global__ void k_cell_count(Fields F, const int32_t* __restrict__ inst_of, const Inst* __restrict__ inst,
                             const int32_t* __restrict__ prev_perm, uint32_t* __restrict__ keys,
                             int32_t* __restrict__ ccount, int64_t n, const Ctl* ctl) {
    if (ctl->abort) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    uint32_t key = 0xFFFFFFFFu;
    if (valid) {
        const int32_t i = prev_perm[t];
        const Inst& s = inst[inst_of[i]];
        const int ix = cell_coord(F.f[SPHMHD_F_X][i], s.lo[0], s.ceff[0], s.n[0]);
        const int iy = cell_coord(F.f[SPHMHD_F_Y][i], s.lo[1], s.ceff[1], s.n[1]);
        const int iz = s.dim == 3 ? cell_coord(F.f[SPHMHD_F_Z][i], s.lo[2], s.ceff[2], s.n[2]) : 0;
        const int64_t local = ix + (int64_t)s.n[0] * (iy + (int64_t)s.n[1] * iz);
        key = (uint32_t)(s.cell_base + local);
        keys[i] = key;
    }
    const unsigned long long heads = run_heads(key, valid);
    const unsigned long long vmask = __ballot(valid);
    const int lane = threadIdx.x & (kWave - 1);
    if ((heads >> lane) & 1ull) atomicAdd(&ccount[key], run_end(heads, vmask, lane) - lane);
}

// exclusive scan of a block-sized value per thread (kBlock threads)
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int& total) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    int incl = v;
    for (int off = 1; off < kWave; off <<= 1) {
        const int o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int base = 0;
    total = 0;
    for (int q = 0; q < kBlock / kWave; ++q) {
        if (q < w) base += wsum[q];
        total += wsum[q];
    }
    return base + incl - v;
}

// exclusive scan of the block totals (one workgroup)
__device__ void cell_scan_top(int32_t* __restrict__ cblock, const Ctl* ctl) {
    __shared__ int wsum[kBlock / kWave];
    const int64_t nb = (ctl->total_cells + kScanCells - 1) / kScanCells;
    const int64_t per = (nb + kBlock - 1) / kBlock;
    const int64_t a = (int64_t)threadIdx.x * per, b = a + per < nb ? a + per : nb;
    int sum = 0;
    for (int64_t q = a; q < b; ++q) sum += ld_dev(&cblock[q]);  // written by this launch
    int tot;
    int run = block_excl_scan(sum, wsum, tot);
    for (int64_t q = a; q < b; ++q) {
        const int c = ld_dev(&cblock[q]);
        cblock[q] = run;
        run += c;
    }
}

// per chunk of kScanCells cells: local exclusive/inclusive prefix into
// cstart/cend and the chunk total into cblock; workgroups stride over the
// chunks of this sort's cells (the launch is sized for the particles, not
// for the capacity); the last workgroup to finish scans the chunk totals
__global__ void __launch_bounds__(kBlock) k_cell_scan(const int32_t* __restrict__ ccount,
                                                      int32_t* __restrict__ cstart, int32_t* __restrict__ cend,
                                                      int32_t* __restrict__ cblock, Ctl* ctl, int32_t* arrive) {
    if (ctl->abort) return;
    __shared__ int wsum[kBlock / kWave];
    const int64_t total = ctl->total_cells;
    for (int64_t chunk = blockIdx.x; chunk * kScanCells < total; chunk += gridDim.x) {  // uniform per workgroup
        const int64_t c0 = chunk * kScanCells;
        const int64_t cb = c0 + (int64_t)threadIdx.x * kScanItems;
        int v[kScanItems];
        int sum = 0;
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            v[q] = cb + q < total ? ccount[cb + q] : 0;
            sum += v[q];
        }
        int btot;
        int run = block_excl_scan(sum, wsum, btot);
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            if (cb + q < total) {
                cstart[cb + q] = run;
                cend[cb + q] = run + v[q];
            }
            run += v[q];
        }
        if (threadIdx.x == 0) st_dev(&cblock[chunk], btot);  // read by the tail
        __syncthreads();  // wsum is reused by the next chunk
    }
    if (last_to_arrive(arrive)) cell_scan_top(cblock, ctl);
}

// each run takes a block of its cell's slots (filled from the end down, so
// the counts return to zero); the order inside a bucket is arbitrary here
__global__ void k_cell_place(const int32_t* __restrict__ prev_perm, const uint32_t* __restrict__ keys,
                             int32_t* __restrict__ ccount, const int32_t* __restrict__ cstart,
                             const int32_t* __restrict__ cblock, int32_t* __restrict__ bucket, int64_t n,
                             const Ctl* ctl) {
    if (ctl->abort) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    const int32_t i = valid ? prev_perm[t] : 0;
    const uint32_t key = valid ? keys[i] : 0xFFFFFFFFu;
    const unsigned long long heads = run_heads(key, valid);
    const unsigned long long vmask = __ballot(valid);
    const int lane = threadIdx.x & (kWave - 1);
    int32_t top = 0;
    if ((heads >> lane) & 1ull) {
        const int len = run_end(heads, vmask, lane) - lane;
        top = cstart[key] + cblock[key / kScanCells] + atomicSub(&ccount[key], len) - len;
    }
    // this lane's run head: the highest head at or below it
    const unsigned long long upto = heads & (lane == kWave - 1 ? ~0ull : ((1ull << (lane + 1)) - 1));
    const int hl = upto ? kWave - 1 - __clzll(upto) : lane;
    const int32_t base = __shfl(top, hl);
    if (valid) bucket[base + (lane - hl)] = i;
}

// the density pack {x, y, z, m} of storage index i
__device__ __forceinline__ double4 density_record(const Fields& F, int32_t i, int dim) {
    double4 r;
    r.x = F.f[SPHMHD_F_X][i];
    r.y = F.f[SPHMHD_F_Y][i];
    r.z = dim == 3 ? F.f[SPHMHD_F_Z][i] : 0.0;
    r.w = F.f[SPHMHD_F_M][i];
    return r;
}

// One launch for the cell table and the small buckets.  Thread q finishes
// cells q, q + threads, ... (start and end in the sorted order from the
// scan's chunk-local prefixes plus the chunk's offset; cells above
// kSmallBucket particles are listed for k_cell_big; their largest-h slots
// cleared) and places the particle in bucket slot q: in a small cell its
// place is its rank by storage index, and with dpack its density record goes
// to that place too (k_pack_density's work, without a launch of its own).
// Both read the scan's prefixes (cls, cle) and write the final table
// (cstart, cend), so neither waits on the other.
__global__ void k_cell_finish_rank(const int32_t* __restrict__ cls, const int32_t* __restrict__ cle,
                                   const int32_t* __restrict__ cblock, int32_t* __restrict__ cstart,
                                   int32_t* __restrict__ cend, int32_t* __restrict__ bigc,
                                   unsigned long long* __restrict__ hcell, const int32_t* __restrict__ bucket,
                                   const uint32_t* __restrict__ keys, int32_t* __restrict__ perm,
                                   uint32_t* __restrict__ skey, int64_t n, int64_t cap, Ctl* ctl, Fields F,
                                   double4* __restrict__ dpack, int dim) {
    if (ctl->abort) return;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t ncell = min(cap, ctl->total_cells);
    for (int64_t c = q; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t off = cblock[c / kScanCells];
        const int32_t s = cls[c] + off, e = cle[c] + off;
        cstart[c] = s;
        cend[c] = e;
        hcell[c] = 0ull;  // filled from the final h before pruned lists are built
        hcell[cap + c] = 0ull;  // the row maxima, at each row's first cell
        if (e - s > kSmallBucket) bigc[atomicAdd(&ctl->nbig, 1)] = (int32_t)c;
    }
    if (q >= n) return;
    const int32_t v = bucket[q];
    const uint32_t c = keys[v];
    const int32_t off = cblock[c / kScanCells];
    const int32_t s = cls[c] + off, e = cle[c] + off;
    if (e - s > kSmallBucket) return;
    int32_t r = 0;
    for (int32_t k = s; k < e; ++k) r += bucket[k] < v ? 1 : 0;
    perm[s + r] = v;
    skey[s + r] = c;
    if (dpack) dpack[s + r] = density_record(F, v, dim);
}

// ascending bitonic network over P = 2^k slots of which the first m are
// real; comparisons with a slot >= m are skipped, which is exact because a
// missing slot stands for +infinity (every compare-exchange puts the smaller
// value at the lower index)
__device__ __forceinline__ void bitonic_sort(int32_t* a, int64_t m, int64_t P) {
    for (int64_t k = 2; k <= P; k <<= 1) {
        for (int64_t j = k >> 1; j > 0; j >>= 1) {
            for (int64_t q = threadIdx.x; q < P; q += blockDim.x) {
                const int64_t p = j == (k >> 1) ? (q ^ (k - 1)) : (q ^ j);
                if (p > q && p < m) {
                    const int32_t x = a[q], y = a[p];
                    if (x > y) {
                        a[q] = y;
                        a[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_cell_big(const int32_t* __restrict__ cstart,
                                                     const int32_t* __restrict__ cend,
                                                     const int32_t* __restrict__ bucket, int32_t* perm,
                                                     uint32_t* __restrict__ skey, const int32_t* __restrict__ bigc,
                                                     const Ctl* ctl, Fields F, double4* __restrict__ dpack, int dim) {
    if (ctl->abort) return;
    __shared__ int32_t lds[kBigLds];
    const int nbig = ctl->nbig;
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int32_t c = bigc[b];
        const int32_t s = cstart[c];
        const int64_t m = cend[c] - s;
        int64_t P = 1;
        while (P < m) P <<= 1;
        for (int64_t q = threadIdx.x; q < m; q += blockDim.x) skey[s + q] = (uint32_t)c;
        if (m <= kBigLds) {
            for (int64_t q = threadIdx.x; q < P; q += blockDim.x) lds[q] = q < m ? bucket[s + q] : INT_MAX;
            __syncthreads();
            bitonic_sort(lds, P, P);
            for (int64_t q = threadIdx.x; q < m; q += blockDim.x) perm[s + q] = lds[q];
        } else {
            for (int64_t q = threadIdx.x; q < m; q += blockDim.x) perm[s + q] = bucket[s + q];
            __syncthreads();
            bitonic_sort(perm + s, m, P);
        }
        __syncthreads();
        if (dpack)
            for (int64_t q = threadIdx.x; q < m; q += blockDim.x) dpack[s + q] = density_record(F, perm[s + q], dim);
    }
}

// ---------------------------------------------------------------------------
// packs
// ---------------------------------------------------------------------------

__global__ void k_pack_density(Fields F, const int32_t* __restrict__ perm, double4* __restrict__ dpack, int64_t n,
                               int dim, const Ctl* ctl) {
    if (ctl->abort) return;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    dpack[t] = density_record(F, perm[t], dim);
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

// split grids: each cell's largest h from the cell-ordered {x,y,z,h} pack,
// which the pruned list build passes cells over by (hcell was cleared by the
// sort's k_cell_finish_rank)
__global__ void k_cell_h(const double4* __restrict__ xh, const uint32_t* __restrict__ skey,
                         const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of,
                         unsigned long long* __restrict__ hcell, int64_t cap, int64_t n, const Ctl* ctl) {
    if (ctl->abort) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const Inst* s = t < n ? &inst[inst_of[t]] : nullptr;
    const bool use = s && s->split > 1;
    const uint32_t key = use ? skey[t] : 0xFFFFFFFFu;
    const double h = use ? xh[t].w : 0.0;
    cell_hmax(hcell, key, h, use);
    // the row's first cell: sorted keys give sorted row keys, so equal rows
    // stay runs of the wave
    uint32_t rkey = key;
    if (use) {
        const int64_t local = (int64_t)key - s->cell_base;
        rkey = (uint32_t)(s->cell_base + local - local % s->n[0]);
    }
    cell_hmax(hcell + cap, rkey, h, use);
}

// One axis of the dilation of the per-cell h maxima over the h_max stencil:
// dst[c] = the largest src over the cells within srad[axis] of c along the
// axis (wrapped on periodic axes, clipped on open ones; the whole axis where
// the stencil spans it).  Instances without per-cell values (not split) get 0.
__global__ void k_cell_dilate(const unsigned long long* __restrict__ src, unsigned long long* __restrict__ dst,
                              const Inst* __restrict__ inst, int ninst, int axis, int64_t cap, const Ctl* ctl) {
    if (ctl->abort) return;
    const int64_t ncell = min(cap, ctl->total_cells);
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = ninst - 1;  // the instance whose cells hold c
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (inst[mid].cell_base <= c) lo = mid;
            else hi = mid - 1;
        }
        const Inst& s = inst[lo];
        if (s.split <= 1) {
            dst[c] = 0ull;
            continue;
        }
        const int64_t local = c - s.cell_base;
        const int n0 = s.n[0], n1 = s.n[1];
        const int co = axis == 0 ? (int)(local % n0) : (axis == 1 ? (int)((local / n0) % n1) : (int)(local / ((int64_t)n0 * n1)));
        const int64_t step = axis == 0 ? 1 : (axis == 1 ? (int64_t)n0 : (int64_t)n0 * n1);
        const int na = s.n[axis], sr = s.srad[axis];
        const int64_t first = c - (int64_t)co * step;  // the cell of coordinate 0 on this line
        unsigned long long m = 0ull;
        if (2 * sr + 1 >= na) {
            for (int q = 0; q < na; ++q) {
                const unsigned long long v = src[first + (int64_t)q * step];
                m = v > m ? v : m;
            }
        } else {
            for (int q = co - sr; q <= co + sr; ++q) {
                int w = q;
                if (w < 0 || w >= na) {
                    if (!s.periodic) continue;
                    w = w < 0 ? w + na : w - na;
                }
                const unsigned long long v = src[first + (int64_t)w * step];
                m = v > m ? v : m;
            }
        }
        dst[c] = m;
    }
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

// one tile of B sorted particles starting at t0
template <int DIM, bool PRUNE>
__device__ __forceinline__ void nb_block(const double4* __restrict__ xh, const uint32_t* __restrict__ skey,
                                         const int32_t* __restrict__ cstart, const int32_t* __restrict__ cend,
                                         const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of,
                                         int32_t* __restrict__ nb, int32_t* __restrict__ nbcnt, int64_t n,
                                         int stride, Ctl* ctl, int cap, const unsigned long long* __restrict__ hcell,
                                         const unsigned long long* __restrict__ hrow,
                                         const unsigned long long* __restrict__ hreach, int64_t t0) {
    // 2-D blocks stage their candidates in LDS (measured 7-20 % faster): cap
    // records of dynamic LDS, sized with the workgroup; in 3-D the staged runs
    // rarely fit and the LDS would cost occupancy (cap 0)
    __shared__ StageHdr H;
    extern __shared__ double4 stage[];
    const int B = blockDim.x;  // 64, 128 or 256 (the launch shape is tuned)
    if (t0 >= n) return;  // whole block past the end (the block barriers below need every wave)
    if constexpr (DIM == 2) {
        if (threadIdx.x < kWave) stage_plan<DIM, kWave>(H, skey, cstart, cend, inst, inst_of, t0, B, n, cap);
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
        // one loop with the ranges' image shifts, one with min_image per
        // candidate (small periodic grids); both give the same r^2 bits
        // PRUNE (chosen where some grid is split: the pruning state costs
        // occupancy): with range shifts the stencil is pruned by cell against
        // h_max, or on split grids against each cell's largest h (hcell, when
        // the caller filled it from the current h)
        // the particle's reach on split grids (hreach: the largest h within
        // the h_max stencil of its cell): every j with r < R max(h_i, h_j)
        // has h_j <= that h, so lies within ceil(R max(h_i, h) / c_eff) cells
        // (an axis scanned whole keeps its order), and the stencil is pruned
        // geometrically against that h, without reading any cell's own
        const bool narrow = PRUNE && hreach && s.split > 1;
        const double hreach_i = narrow ? bitsd(hreach[skey[t]]) : 0.0;
        Prune pr;
        if (PRUNE && s.range_shift) {
            // a particle whose reach passes its own h (near larger h) drops
            // and trims each row by the row's largest h: only rows holding
            // such h need the long reach (identical lists, fewer candidates)
            if (narrow)
                prune_init<DIM>(pr, s, (int64_t)skey[t], xi, yi, zi, R, hi, hreach_i, nullptr,
                                hrow && hreach_i > hi ? hrow : nullptr);
            else
                prune_init<DIM>(pr, s, (int64_t)skey[t], xi, yi, zi, R, hi, bitsd(s.hmax_bits),
                                hcell && s.split > 1 ? hcell : nullptr, hcell && s.split > 1 ? hrow : nullptr);
        }
        int sri[3] = {s.srad[0], s.srad[1], s.srad[2]};
        if (narrow) {
            const double reach = R * fmax(hi, hreach_i);
            for (int a = 0; a < DIM; ++a) {
                if (2 * s.srad[a] + 1 >= s.n[a]) continue;
                const double q = ceil(reach / s.ceff[a]);
                sri[a] = q < 1.0 ? 1 : (q < (double)s.srad[a] ? (int)q : s.srad[a]);
            }
        }
        auto scan = [&](auto shift_c) {
            constexpr bool kShift = decltype(shift_c)::value;
            for_each_candidate<DIM>(s, (int64_t)skey[t], cstart, cend, H, stage, xh,
                                    [&](int j, const double4& o, double sx, double sy, double sz) {
                const double dx = kShift ? (xi - o.x) + sx : min_image(xi - o.x, L0, H0, per);
                const double dy = kShift ? (yi - o.y) + sy : min_image(yi - o.y, L1, H1, per);
                const double dz = DIM == 3 ? (kShift ? (zi - o.z) + sz : min_image(zi - o.z, L2, H2, per)) : 0.0;
                const double r2 = r2_exact(dx, dy, dz);
                const double rad = R * fmax(hi, o.w);
                if (r2 < rad * rad) {
                    if (cnt < stride) nb[(int64_t)cnt * n + t] = j;
                    ++cnt;
                }
            }, PRUNE && kShift ? &pr : nullptr, narrow ? sri : nullptr);
        };
        if (s.range_shift) scan(std::true_type{});
        else scan(std::false_type{});
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
        for (int w = 1; w < B / kWave; ++w) {
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

template <int DIM, bool PRUNE>
__global__ void __launch_bounds__(kBlock) k_nb_build(const double4* __restrict__ xh, const uint32_t* __restrict__ skey,
                                                     const int32_t* __restrict__ cstart, const int32_t* __restrict__ cend,
                                                     const Inst* __restrict__ inst, const int32_t* __restrict__ inst_of,
                                                     int32_t* __restrict__ nb, int32_t* __restrict__ nbcnt, int64_t n,
                                                     int stride, Ctl* ctl, int cap,
                                                     const unsigned long long* __restrict__ hcell,
                                                     const unsigned long long* __restrict__ hrow,
                                                     const unsigned long long* __restrict__ hreach, int32_t* queue,
                                                     int32_t* arrive) {
    if (ctl->abort) return;
    const int B = blockDim.x;
    if (queue) {  // persistent: tiles from the work queues, then the last block resets them
        int polls = 0;
        for (int64_t tile; (tile = pull_tile(queue, (n + B - 1) / B, polls)) >= 0;)
            nb_block<DIM, PRUNE>(xh, skey, cstart, cend, inst, inst_of, nb, nbcnt, n, stride, ctl, cap, hcell,
                                 hrow, hreach, tile * B);
        if (last_to_arrive(arrive)) reset_queues(queue);
    } else {
        nb_block<DIM, PRUNE>(xh, skey, cstart, cend, inst, inst_of, nb, nbcnt, n, stride, ctl, cap, hcell,
                             hrow, hreach, xcd_tile((n + B - 1) / B) * B);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void launch_eval_begin(Dev& d, bool reset_dt) {
    hipLaunchKernelGGL(k_eval_begin, dim3(1), dim3(kBlock), 0, d.stream, d.inst, d.ninst, d.ctl, reset_dt ? 1 : 0);
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

// h_max per instance, then (the launch's last workgroup) the cell grid
void launch_hmax_grid(Dev& d) {
    if (d.n == 0) {
        hipLaunchKernelGGL(k_grid, dim3(1), dim3(kBlock), 0, d.stream, d.inst, d.ninst, d.ctl);
        return;
    }
    const int64_t blocks = std::min<int64_t>(kHmaxBlocks, nblk(d.n));
    const int64_t chunk = (d.n + blocks - 1) / blocks;
    hipLaunchKernelGGL(k_hmax, dim3((unsigned)blocks), dim3(kBlock), 0, d.stream, d.fields, d.inst_of, d.inst, d.n,
                       chunk, d.ninst, d.ctl, d.arrive);
}

constexpr int kBigBlocks = 256;

void launch_keys_sort_ranges(Dev& d, bool pack) {
    if (d.n == 0) return;
    // keys_alt: keys by storage index; perm_alt: the buckets; perm: the
    // previous order on entry, the new one on exit
    hipLaunchKernelGGL(k_cell_count, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.inst_of, d.inst, d.perm,
                       d.keys_alt, d.ccount, d.n, d.ctl);
    // chunks of kScanCells cells: as many workgroups as the particles would
    // fill (cells beyond are strided over), at most the capacity's chunks
    const int64_t chunks = (d.cell_cap + kScanCells - 1) / kScanCells;
    const unsigned sblocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(chunks, (d.n + kScanCells - 1) / kScanCells));
    hipLaunchKernelGGL(k_cell_scan, dim3(sblocks), dim3(kBlock), 0, d.stream, d.ccount, d.cls, d.cle, d.cblock,
                       d.ctl, d.arrive);
    hipLaunchKernelGGL(k_cell_place, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.perm, d.keys_alt, d.ccount,
                       d.cls, d.cblock, d.perm_alt, d.n, d.ctl);
    double4* dp = pack ? (double4*)d.dpack : nullptr;
    hipLaunchKernelGGL(k_cell_finish_rank, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.cls, d.cle, d.cblock,
                       d.cstart, d.cend, d.bigc, d.hcell, d.perm_alt, d.keys_alt, d.perm, d.keys, d.n, d.cell_cap,
                       d.ctl, d.fields, dp, d.dim);
    hipLaunchKernelGGL(k_cell_big, dim3(kBigBlocks), dim3(kBlock), 0, d.stream, d.cstart, d.cend, d.perm_alt, d.perm,
                       d.keys, d.bigc, d.ctl, d.fields, dp, d.dim);
}

void launch_pack_density(Dev& d) {
    if (d.n == 0) return;
    hipLaunchKernelGGL(k_pack_density, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, d.fields, d.perm,
                       (double4*)d.dpack, d.n, d.dim, d.ctl);
}

template <int DIM, bool PRUNE>
static void nb_build(Dev& d, const unsigned long long* hc, const unsigned long long* hreach) {
    const int B = d.blk[LC_LISTS];
    // 2-D blocks stage candidate runs in LDS, cap records sized with the block
    const int cap = DIM == 2 ? kStageCap * B / kBlock : 0;
    int32_t* q = d.persist[LC_LISTS] ? d.queue : nullptr;
    const unsigned grid = q ? persistent_grid((const void*)k_nb_build<DIM, PRUNE>, B, sizeof(double4) * cap,
                                              (d.n + B - 1) / B)
                            : xcd_grid((d.n + B - 1) / B);
    hipLaunchKernelGGL((k_nb_build<DIM, PRUNE>), dim3(grid), dim3(B), sizeof(double4) * cap, d.stream,
                       (const double4*)d.dpack, d.keys, d.cstart, d.cend, d.inst, d.inst_of, d.nb, d.nbcnt, d.n,
                       d.nb_stride, d.ctl, cap, hc, hc ? hc + d.cell_cap : nullptr, hreach, q, d.arrive);
}

void launch_nb_build(Dev& d, bool cell_h) {
    if (d.n == 0) return;
    const bool per_cell = cell_h && d.prune_lists;
    if (per_cell)
        hipLaunchKernelGGL(k_cell_h, dim3(nblk(d.n)), dim3(kBlock), 0, d.stream, (const double4*)d.dpack, d.keys,
                           d.inst, d.inst_of, d.hcell, d.cell_cap, d.n, d.ctl);
    const unsigned long long* hc = per_cell ? d.hcell : nullptr;
    // 3-D: the largest h within each cell's h_max stencil (the cell maxima
    // dilated along x, y, z), which bounds each particle's reach
    const unsigned long long* hreach = nullptr;
    if (per_cell && d.dim == 3) {
        unsigned long long* t1 = d.hcell + 2 * d.cell_cap;
        unsigned long long* t2 = d.hcell + 3 * d.cell_cap;
        const unsigned g = nblk(d.n);
        hipLaunchKernelGGL(k_cell_dilate, dim3(g), dim3(kBlock), 0, d.stream, (const unsigned long long*)d.hcell, t1,
                           d.inst, d.ninst, 0, d.cell_cap, d.ctl);
        hipLaunchKernelGGL(k_cell_dilate, dim3(g), dim3(kBlock), 0, d.stream, (const unsigned long long*)t1, t2,
                           d.inst, d.ninst, 1, d.cell_cap, d.ctl);
        hipLaunchKernelGGL(k_cell_dilate, dim3(g), dim3(kBlock), 0, d.stream, (const unsigned long long*)t2, t1,
                           d.inst, d.ninst, 2, d.cell_cap, d.ctl);
        hreach = t1;
    }
    if (d.dim == 2) {
        if (d.prune_lists) nb_build<2, true>(d, hc, nullptr);
        else nb_build<2, false>(d, hc, nullptr);
    } else {
        if (d.prune_lists) nb_build<3, true>(d, hc, hreach);
        else nb_build<3, false>(d, hc, nullptr);
    }
}

}  // namespace sph
