// sort.hip -- spatial order of the partner array: a cell list as an integer key sort plus a
// prefix scan.
//
// Every particle gets a 32-bit key: its class in the top two bits (0 CG-DNA site, 1 fixed
// ion, 2 moving ion) over the Morton code of its cell on a regular lattice over the box
// (10 bits per axis), so each class is one run of the order and its tiles hold one class
// only, but for the tile a run boundary falls in.  A stable least-significant-digit radix sort of (key, particle index)
// -- four passes of 8 bits, each a per-block digit histogram, one exclusive scan of the
// histograms in (digit, block) order and a stable scatter -- gives the partner order; a
// second scan over the sorted keys numbers the occupied cells and records the first slot of
// each (cell_start).  A gather writes the partner array spos (x, y, z, packed type and
// original index) in that order plus rank[i], the slot of particle i.  The step kernels
// keep spos current between sorts by writing each updated particle to its slot, so a sort
// is only needed every few steps; the order only decides which partners share a wavefront
// or a tile, never which pairs are summed.
#include <utility>

#include "engine.h"

namespace bromoc {

constexpr int SORT_BLOCK = 256;        // threads of the histogram / scatter blocks
constexpr int SORT_RADIX = 256;        // 8-bit digits
constexpr int SORT_PASSES = 4;         // 32 bits of key (class 2, Morton code 30)
constexpr int SCAN_BLOCK = 1024;
constexpr int SWAVE = 64;              // lanes of a wavefront

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t cell_coord(double x, double lo, double inv, int n) {
    double c = (x - lo) * inv;
    int k = c > 0.0 ? (int)c : 0;   // also catches NaN
    return (uint32_t)(k < n ? k : n - 1);
}

__device__ __forceinline__ uint32_t sort_key(const double4 &p, int i, const SortGeom &g) {
    const uint32_t cx = cell_coord(p.x, g.lo[0], g.inv, g.n[0]);
    const uint32_t cy = cell_coord(p.y, g.lo[1], g.inv, g.n[1]);
    const uint32_t cz = cell_coord(p.z, g.lo[2], g.inv, g.n[2]);
    const uint32_t cls = i < g.nsites ? 0u : i < g.nsites + g.nfix ? 1u : 2u;
    return (cls << 30) | spread10(cx) | (spread10(cy) << 1) | (spread10(cz) << 2);
}

__global__ void k_keys(const double4 *__restrict__ pos, int n, SortGeom g, uint32_t *keys,
                       int32_t *vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = sort_key(pos[i], i, g);
    vals[i] = i;
}

// digit counts of block b's chunk [b*chunk, min((b+1)*chunk, n)): hist[d * nb + b]
__global__ __launch_bounds__(SORT_BLOCK) void k_digit_hist(const uint32_t *__restrict__ keys,
                                                            int n, int chunk, int shift,
                                                            int32_t *hist) {
    __shared__ int32_t s_c[SORT_RADIX];
    const int b = blockIdx.x, nb = gridDim.x, tid = threadIdx.x;
    for (int d = tid; d < SORT_RADIX; d += SORT_BLOCK) s_c[d] = 0;
    __syncthreads();
    const int lo = b * chunk, hi = min(lo + chunk, n);
    for (int k = lo + tid; k < hi; k += SORT_BLOCK)
        atomicAdd(&s_c[(keys[k] >> shift) & (SORT_RADIX - 1)], 1);   // integer: exact
    __syncthreads();
    for (int d = tid; d < SORT_RADIX; d += SORT_BLOCK) hist[d * nb + b] = s_c[d];
}

// exclusive prefix sum of a[0..m) in place, one block, in rounds of SCAN_BLOCK * SCAN_PER
// elements: thread t loads its SCAN_PER consecutive elements of the round at once (16-byte
// loads, all in flight together), the thread totals are scanned by wavefront shuffles and
// the wavefront totals in LDS, and each thread writes its elements' prefixes; a running
// carry joins the rounds.  Also total[0] = the sum.  Integer sums: exact in any order.
constexpr int SCAN_PER = 64;
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_excl(int32_t *a, int m, int32_t *total) {
    constexpr int NWS = SCAN_BLOCK / SWAVE, NV = SCAN_PER / 4;
    __shared__ int32_t s_w[NWS];
    __shared__ int32_t s_carry;
    const int tid = threadIdx.x, lane = tid & (SWAVE - 1), wv = tid / SWAVE;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < m; base += SCAN_BLOCK * SCAN_PER) {
        const int lo = base + tid * SCAN_PER;
        int4 v[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) {
            const int k = lo + 4 * q;
            if (k + 3 < m) {
                v[q] = *reinterpret_cast<const int4 *>(a + k);
            } else {
                v[q].x = k < m ? a[k] : 0;
                v[q].y = k + 1 < m ? a[k + 1] : 0;
                v[q].z = k + 2 < m ? a[k + 2] : 0;
                v[q].w = 0;
            }
        }
        int32_t sum = 0;
#pragma unroll
        for (int q = 0; q < NV; q++) sum += v[q].x + v[q].y + v[q].z + v[q].w;
        int32_t incl = sum;
#pragma unroll
        for (int off = 1; off < SWAVE; off <<= 1) {
            const int32_t u = __shfl_up(incl, off, SWAVE);
            if (lane >= off) incl += u;
        }
        if (lane == SWAVE - 1) s_w[wv] = incl;
        __syncthreads();
        int32_t run = s_carry + incl - sum, tot = 0;
        for (int q = 0; q < NWS; q++) {
            if (q < wv) run += s_w[q];
            tot += s_w[q];
        }
#pragma unroll
        for (int q = 0; q < NV; q++) {
            const int k = lo + 4 * q;
            int4 o;
            o.x = run; run += v[q].x;
            o.y = run; run += v[q].y;
            o.z = run; run += v[q].z;
            o.w = run; run += v[q].w;
            if (k + 3 < m) {
                *reinterpret_cast<int4 *>(a + k) = o;
            } else {
                if (k < m) a[k] = o.x;
                if (k + 1 < m) a[k + 1] = o.y;
                if (k + 2 < m) a[k + 2] = o.z;
            }
        }
        __syncthreads();
        if (tid == 0) s_carry += tot;
        __syncthreads();
    }
    if (tid == 0 && total) total[0] = s_carry;
}

// the lanes of this wavefront (among those with ok set) whose 8-bit digit equals d
__device__ __forceinline__ uint64_t digit_peers(bool ok, int d) {
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
        const uint64_t m = __ballot((d >> bit) & 1);
        peers &= ((d >> bit) & 1) ? m : ~m;
    }
    return peers;
}

// stable scatter of block b's chunk: element k of digit d goes to offset[d * nb + b] plus
// the number of earlier elements of the chunk with digit d.  The chunk is taken in tiles of
// SORT_BLOCK in order; within a tile, lanes of equal digit are ranked by lane (eight
// ballots over the digit's bits) and wavefronts in order.
__global__ __launch_bounds__(SORT_BLOCK) void k_digit_scatter(
    const uint32_t *__restrict__ kin, const int32_t *__restrict__ vin, uint32_t *kout,
    int32_t *vout, int n, int chunk, int shift, const int32_t *__restrict__ offset) {
    constexpr int NWV = SORT_BLOCK / 64;
    __shared__ int32_t s_run[SORT_RADIX];
    __shared__ int32_t s_wc[NWV][SORT_RADIX];
    const int b = blockIdx.x, nb = gridDim.x, tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    for (int d = tid; d < SORT_RADIX; d += SORT_BLOCK) {
        s_run[d] = offset[d * nb + b];
        for (int w = 0; w < NWV; w++) s_wc[w][d] = 0;
    }
    __syncthreads();
    const int lo = b * chunk, hi = min(lo + chunk, n);
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int base = lo; base < hi; base += SORT_BLOCK) {
        const int k = base + tid;
        const bool ok = k < hi;
        const uint32_t key = ok ? kin[k] : 0u;
        const int32_t val = ok ? vin[k] : 0;
        const int d = (key >> shift) & (SORT_RADIX - 1);
        const uint64_t peers = digit_peers(ok, d);   // lanes of this wavefront with digit d
        const int rank = __popcll(peers & lt);
        if (ok && rank == 0) s_wc[wv][d] = __popcll(peers);
        __syncthreads();
        if (ok) {
            int before = s_run[d] + rank;
            for (int w = 0; w < wv; w++) before += s_wc[w][d];
            kout[before] = key;
            vout[before] = val;
        }
        __syncthreads();
        for (int dd = tid; dd < SORT_RADIX; dd += SORT_BLOCK) {
            int t = 0;
            for (int w = 0; w < NWV; w++) { t += s_wc[w][dd]; s_wc[w][dd] = 0; }
            s_run[dd] += t;
        }
        __syncthreads();
    }
}

// cell numbering: flag[s] = 1 where the sorted key changes (and at slot 0)
__global__ void k_cell_flags(const uint32_t *__restrict__ keys, int n, int32_t *flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    flag[s] = (s == 0 || keys[s] != keys[s - 1]) ? 1 : 0;
}

// after the exclusive scan of the flags: cell_start[cell of s] = s at every flagged slot,
// and cell_start[ncell] = n
__global__ void k_cell_starts(const uint32_t *__restrict__ keys, const int32_t *__restrict__ excl,
                              int n, const int32_t *ncell, int32_t *cell_start) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) cell_start[ncell[0]] = n;
    if (s >= n) return;
    if (s == 0 || keys[s] != keys[s - 1]) cell_start[excl[s]] = s;
}

__global__ void k_gather(const double4 *__restrict__ pos, const int32_t *__restrict__ order, int n,
                         double4 *spos0, double4 *spos1, int32_t *rank) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = order[k];
    const double4 p = pos[j];
    spos0[k] = p;
    spos1[k] = p;
    rank[j] = k;
}

// The whole spatial order of a small system (n <= SMALL_SORT_MAX) in one block, in LDS: the
// keys, the same four stable 8-bit passes (digit histogram by one LDS atomic per digit and
// wavefront, its exclusive scan by one wavefront, then the ballot-ranked scatter in tiles
// of the block, in order), the sorted keys, order, rank and partner array, and the cell
// starts from a block scan of the key changes.  One launch instead of the nineteen of the
// multi-block path, whose launches dominate at a few thousand particles; the outputs are
// the same.
constexpr int SMALL_SORT_THREADS = 1024;
constexpr int SMALL_SORT_PER = 4;
constexpr int SMALL_SORT_MAX = SMALL_SORT_THREADS * SMALL_SORT_PER;
__global__ __launch_bounds__(SMALL_SORT_THREADS) void k_sort_small(
    const double4 *__restrict__ pos, int n, SortGeom g, uint32_t *keys_out, int32_t *order,
    int32_t *rank, int32_t *cell_start, int32_t *ncell, double4 *spos0, double4 *spos1) {
    constexpr int NWV = SMALL_SORT_THREADS / SWAVE;
    __shared__ uint32_t s_k[2][SMALL_SORT_MAX];
    __shared__ int32_t s_v[2][SMALL_SORT_MAX];
    __shared__ int32_t s_run[SORT_RADIX];
    __shared__ int32_t s_wc[NWV][SORT_RADIX];   // a tile's digit counts per wavefront
    __shared__ int32_t s_wo[NWV][SORT_RADIX];   // and the slots their runs start at
    __shared__ int32_t s_tot[NWV];
    const int tid = threadIdx.x, lane = tid & (SWAVE - 1), wv = tid / SWAVE;
    for (int i = tid; i < n; i += SMALL_SORT_THREADS) {
        s_k[0][i] = sort_key(pos[i], i, g);
        s_v[0][i] = i;
    }
    for (int q = tid; q < NWV * SORT_RADIX; q += SMALL_SORT_THREADS) (&s_wc[0][0])[q] = 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    int cur = 0;
    for (int pass = 0; pass < SORT_PASSES; pass++) {
        const int shift = 8 * pass;
        for (int d = tid; d < SORT_RADIX; d += SMALL_SORT_THREADS) s_run[d] = 0;
        __syncthreads();
        // digit counts: the lanes of a wavefront holding one digit add their count at once
        // (most keys share their high digits, so per-key atomics would queue on one counter)
        for (int base = 0; base < n; base += SMALL_SORT_THREADS) {
            const int k = base + tid;
            const bool ok = k < n;
            const int d = ok ? (s_k[cur][k] >> shift) & (SORT_RADIX - 1) : 0;
            const uint64_t peers = digit_peers(ok, d);
            if (ok && __popcll(peers & lt) == 0) atomicAdd(&s_run[d], __popcll(peers));   // integer: exact
        }
        __syncthreads();
        // a pass whose digit is the same for every key leaves the order as it is (every
        // thread reads the count before any thread changes s_run again)
        const bool same = s_run[(s_k[cur][0] >> shift) & (SORT_RADIX - 1)] == n;
        __syncthreads();
        if (same) continue;
        if (wv == 0) {   // exclusive scan of the 256 digit counts, four per lane
            int c[SORT_RADIX / SWAVE], sum = 0;
#pragma unroll
            for (int q = 0; q < SORT_RADIX / SWAVE; q++) {
                c[q] = s_run[(SORT_RADIX / SWAVE) * lane + q];
                sum += c[q];
            }
            int incl = sum;
#pragma unroll
            for (int off = 1; off < SWAVE; off <<= 1) {
                const int u = __shfl_up(incl, off, SWAVE);
                if (lane >= off) incl += u;
            }
            int run = incl - sum;
#pragma unroll
            for (int q = 0; q < SORT_RADIX / SWAVE; q++) {
                s_run[(SORT_RADIX / SWAVE) * lane + q] = run;
                run += c[q];
            }
        }
        __syncthreads();
        for (int base = 0; base < n; base += SMALL_SORT_THREADS) {
            const int k = base + tid;
            const bool ok = k < n;
            const uint32_t key = ok ? s_k[cur][k] : 0u;
            const int32_t val = ok ? s_v[cur][k] : 0;
            const int d = (key >> shift) & (SORT_RADIX - 1);
            const uint64_t peers = digit_peers(ok, d);
            const int r = __popcll(peers & lt);
            if (ok && r == 0) s_wc[wv][d] = __popcll(peers);
            __syncthreads();
            // per digit, in wavefront order: the slot where each wavefront's run of it
            // starts; the counts are cleared for the next tile
            for (int dd = tid; dd < SORT_RADIX; dd += SMALL_SORT_THREADS) {
                int run = s_run[dd];
                for (int w = 0; w < NWV; w++) {
                    s_wo[w][dd] = run;
                    run += s_wc[w][dd];
                    s_wc[w][dd] = 0;
                }
                s_run[dd] = run;
            }
            __syncthreads();
            if (ok) {
                const int at = s_wo[wv][d] + r;
                s_k[cur ^ 1][at] = key;
                s_v[cur ^ 1][at] = val;
            }
            __syncthreads();
        }
        cur ^= 1;
    }
    // the sorted keys and particles are in s_k[cur], s_v[cur]
    for (int k = tid; k < n; k += SMALL_SORT_THREADS) {
        const int j = s_v[cur][k];
        keys_out[k] = s_k[cur][k];
        order[k] = j;
        rank[j] = k;
        const double4 p = pos[j];
        spos0[k] = p;
        spos1[k] = p;
    }
    // cells: slots where the key changes, numbered by a block scan (thread t: slots 4t..4t+3)
    int f[SMALL_SORT_PER], sum = 0;
#pragma unroll
    for (int q = 0; q < SMALL_SORT_PER; q++) {
        const int sl = SMALL_SORT_PER * tid + q;
        f[q] = (sl < n && (sl == 0 || s_k[cur][sl] != s_k[cur][sl - 1])) ? 1 : 0;
        sum += f[q];
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < SWAVE; off <<= 1) {
        const int u = __shfl_up(incl, off, SWAVE);
        if (lane >= off) incl += u;
    }
    if (lane == SWAVE - 1) s_tot[wv] = incl;
    __syncthreads();
    int run = incl - sum, tot = 0;
    for (int w = 0; w < NWV; w++) {
        if (w < wv) run += s_tot[w];
        tot += s_tot[w];
    }
#pragma unroll
    for (int q = 0; q < SMALL_SORT_PER; q++) {
        if (f[q]) cell_start[run] = SMALL_SORT_PER * tid + q;
        run += f[q];
    }
    if (tid == 0) {
        cell_start[tot] = n;
        ncell[0] = tot;
    }
}

// blocks of the radix passes for n keys: enough chunks to spread the work, few enough
// that the (digit, block) histogram stays small for the one-block scan
static int sort_blocks(int n) {
    const int nb = (n + 4 * SORT_BLOCK - 1) / (4 * SORT_BLOCK);
    return nb < 1 ? 1 : (nb > 128 ? 128 : nb);
}

size_t sort_hist_ints(int n) { return (size_t)SORT_RADIX * sort_blocks(n) + 2; }

hipError_t sort_partners(const SortBufs &b, const double4 *pos, int n, double4 *spos0,
                         double4 *spos1, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n <= SMALL_SORT_MAX) {
        hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(SMALL_SORT_THREADS), 0, s, pos, n, b.geom,
                           b.keys_out, b.order, b.rank, b.cell_start, b.ncell, spos0, spos1);
        return hipGetLastError();
    }
    const int blk = 256, grid = (n + blk - 1) / blk;
    hipLaunchKernelGGL(k_keys, dim3(grid), dim3(blk), 0, s, pos, n, b.geom, b.keys, b.vals);
    const int nb = sort_blocks(n);
    const int chunk = (n + nb - 1) / nb;
    uint32_t *kin = b.keys, *kout = b.keys_out;
    int32_t *vin = b.vals, *vout = b.order;
    for (int pass = 0; pass < SORT_PASSES; pass++) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_digit_hist, dim3(nb), dim3(SORT_BLOCK), 0, s, kin, n, chunk, shift,
                           b.hist);
        hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(SCAN_BLOCK), 0, s, b.hist,
                           SORT_RADIX * nb, (int32_t *)nullptr);
        hipLaunchKernelGGL(k_digit_scatter, dim3(nb), dim3(SORT_BLOCK), 0, s, kin, vin, kout, vout,
                           n, chunk, shift, b.hist);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    // an even number of passes leaves the sorted keys and order in keys / vals; order
    // and keys_out take them for the gather and the cells
    hipLaunchKernelGGL(k_cell_flags, dim3(grid), dim3(blk), 0, s, kin, n, b.cflag);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(SCAN_BLOCK), 0, s, b.cflag, n, b.ncell);
    hipLaunchKernelGGL(k_cell_starts, dim3(grid), dim3(blk), 0, s, kin, b.cflag, n, b.ncell,
                       b.cell_start);
    (void)hipMemcpyAsync(b.order, vin, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(b.keys_out, kin, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
    hipLaunchKernelGGL(k_gather, dim3(grid), dim3(blk), 0, s, pos, b.order, n, spos0, spos1,
                       b.rank);
    return hipGetLastError();
}

}  // namespace bromoc
