// sort.hip — the scans and sorts of the substep, hand-written for gfx950.
//
//   * exclusive_sum_u32 / exclusive_sum_over_cap: a three-phase scan.  Each 4096-element
//     tile is summed, one workgroup scans the tile sums, and each tile is then scanned
//     through LDS (coalesced loads and stores, a blocked scan per thread, a wavefront
//     scan of the thread totals) and offset by its tile's prefix.  Integer sums, so the
//     result does not depend on the launch shape.
//   * the particle sort by cell (classify, P2G and G2P read cell segments) is a counting
//     sort.  Particles arrive close to cell order (they were sorted a substep ago and
//     move at most a cell), so the histogram and the scatter take one atomic per run of
//     equal cells in a wavefront.  Runs of one cell from different wavefronts land in
//     atomic order, so each segment is then put in the order of its order key: the
//     reference rank in exact mode (segments in reference order), the current index
//     otherwise (a stable sort).  Keys are unique, so the order does not depend on how
//     the runs landed.  A segment of up to 16 entries is sorted by one thread in
//     registers (a bitonic network); a longer one by a workgroup: bitonic runs in LDS,
//     then merge passes.
//   * sort_keys_u32 / sort_pairs_u32 (culled ranks; the single-thread redistribution
//     path's active particles): bitonic runs in LDS, then merge passes over the array.
#include <hip/hip_runtime.h>

#include <utility>

#include "host.hpp"
#include "sort.hpp"

// Returns a HIP error from the enclosing function.
#define SDR_RET(x)                 \
  do {                             \
    const hipError_t e_ = (x);     \
    if (e_ != hipSuccess) return e_; \
  } while (0)

namespace sdr {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096
constexpr int kSumThreads = 1024;                     // the tile-sum scan workgroup

constexpr int kRunThreads = 256;
constexpr uint32_t kRun = 2048;  // entries a workgroup sorts in LDS (16 KB of 64-bit keys)
constexpr uint32_t kRegSort = 16;  // longest segment one thread sorts in registers
constexpr uint32_t kNoCell = 0xffffffffu;

inline unsigned blocks_for(int64_t n, int block) {
  const int64_t b = (n + block - 1) / block;
  return (unsigned)(b < 1 ? 1 : b);
}

struct Ident {
  __device__ uint32_t operator()(uint32_t v) const { return v; }
};
struct OverCap {
  uint32_t cap;
  int sizes;
  __device__ uint32_t operator()(uint32_t c) const { return c > cap ? (sizes ? c : 1u) : 0u; }
};

// LDS index of tile element x, padded one word per 32 so the blocked reads (stride 16)
// spread over the banks.
__device__ inline int pad(int x) { return x + (x >> 5); }

// Inclusive scan over the lanes of a wavefront.
__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive scan of one value per thread over the workgroup; *total gets the sum.
template <int NT>
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[NT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v, lane);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const uint32_t s = wsum[w];
    before += w < wave ? s : 0u;
    tot += s;
  }
  __syncthreads();  // wsum may be reused by the caller's next scan
  *total = tot;
  return before + inc - v;
}

// Phase 1: the sum of each tile.
template <class F>
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const uint32_t* __restrict__ in, int64_t n, F f,
                                                             uint32_t* __restrict__ tsum) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t e = base + q * kScanThreads + threadIdx.x;
    if (e < n) s += f(in[e]);
  }
  uint32_t tot;
  (void)block_excl_scan<kScanThreads>(s, &tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// Phase 2: the tile sums become their exclusive prefixes (one workgroup; each thread
// owns a run of consecutive tiles).
__global__ void __launch_bounds__(kSumThreads) k_scan_tsum(uint32_t* __restrict__ tsum, int64_t nt) {
  const int64_t per = (nt + kSumThreads - 1) / kSumThreads;
  const int64_t b0 = (int64_t)threadIdx.x * per;
  const int64_t b1 = b0 + per < nt ? b0 + per : nt;
  uint32_t s = 0;
  for (int64_t b = b0; b < b1; ++b) s += tsum[b];
  uint32_t tot;
  uint32_t run = block_excl_scan<kSumThreads>(s, &tot);
  for (int64_t b = b0; b < b1; ++b) {
    const uint32_t v = tsum[b];
    tsum[b] = run;
    run += v;
  }
}

// Phase 3: each tile scanned and offset by its prefix.  INCL: out_incl also receives
// the inclusive sums (may alias `in`: every element is read before it is written, by
// the same thread).
template <class F, bool INCL>
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* in, int64_t n, F f,
                                                             const uint32_t* __restrict__ tsum,
                                                             uint32_t* __restrict__ out, uint32_t* out_incl) {
  __shared__ uint32_t buf[kScanTile + kScanTile / 32];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int t = threadIdx.x;
  uint32_t v[kScanItems];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {  // coalesced: element q * 256 + t of the tile
    const int64_t e = base + q * kScanThreads + t;
    v[q] = e < n ? f(in[e]) : 0u;
    buf[pad(q * kScanThreads + t)] = v[q];
  }
  __syncthreads();
  uint32_t loc[kScanItems];  // blocked: elements t * 16 .. t * 16 + 15
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    loc[q] = s;
    s += buf[pad(t * kScanItems + q)];
  }
  uint32_t tot;
  const uint32_t off = block_excl_scan<kScanThreads>(s, &tot) + tsum[blockIdx.x];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) buf[pad(t * kScanItems + q)] = loc[q] + off;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const int64_t e = base + q * kScanThreads + t;
    if (e < n) {
      const uint32_t x = buf[pad(q * kScanThreads + t)];
      out[e] = x;
      if (INCL) out_incl[e] = x + v[q];
    }
  }
}

template <class F, bool INCL>
hipError_t scan_launch(void* temp, size_t temp_bytes, const uint32_t* in, F f, uint32_t* out, uint32_t* incl,
                       int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int64_t nt = (n + kScanTile - 1) / kScanTile;
  if (temp_bytes < (size_t)nt * 4) return hipErrorInvalidValue;
  uint32_t* tsum = (uint32_t*)temp;
  k_scan_tiles<F><<<(unsigned)nt, kScanThreads, 0, st>>>(in, n, f, tsum);
  k_scan_tsum<<<1, kSumThreads, 0, st>>>(tsum, nt);
  k_scan_apply<F, INCL><<<(unsigned)nt, kScanThreads, 0, st>>>(in, n, f, tsum, out, incl);
  return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}
void fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st) {
  if (n <= 0) return;
  unsigned b = blocks_for(n, 256);
  if (b > 8192) b = 8192;
  k_fill_u32<<<b, 256, 0, st>>>(p, n, v);
}

// ------------------------------------------------------------------ runs of equal keys
// A lane's run of equal keys in its wavefront (lanes are consecutive particles): the
// run's first lane, and for that first lane the run's length.
struct Run {
  int start;
  uint32_t len;
  bool head;
};
__device__ inline Run run_of(uint32_t c, int lane) {
  const uint32_t prev = __shfl_up(c, 1);
  const bool head = lane == 0 || prev != c;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  Run r;
  r.head = head;
  r.start = 63 - __clzll((long long)(heads & upto));
  const unsigned long long after = heads & ~upto;
  const int end = after ? __ffsll((long long)after) - 1 : 64;
  r.len = head ? (uint32_t)(end - lane) : 0u;
  return r;
}

// Histogram of the particles' clamped cells (every particle scatters through it,
// grid.rs:52-88).
template <class T>
__global__ void k_cell_count(Parts<T> p, int64_t n, GridD<T> g, uint32_t* __restrict__ key,
                             uint32_t* __restrict__ cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t c = kNoCell;
  if (t < n) {
    c = (uint32_t)g.clamped_cell(mk3(p.px[t], p.py[t], p.pz[t]));
    key[t] = c;
  }
  const Run r = run_of(c, lane);
  if (r.head && c != kNoCell) atomicAdd(&cnt[c], r.len);
}

// Each particle takes a slot of its cell's segment; a run takes its slots with one
// atomic, in particle order.
__global__ void k_cell_scatter(const uint32_t* __restrict__ key, int64_t n, const uint32_t* __restrict__ cstart,
                               uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm,
                               const uint32_t* __restrict__ rank, uint32_t* __restrict__ skey) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const uint32_t c = t < n ? key[t] : kNoCell;
  const Run r = run_of(c, lane);
  uint32_t base = 0;
  if (r.head && c != kNoCell) base = atomicAdd(&cursor[c], r.len);
  base = __shfl(base, r.start);
  if (c == kNoCell) return;
  const uint32_t slot = cstart[c] + base + (uint32_t)(lane - r.start);
  perm[slot] = (uint32_t)t;
  if (skey) skey[slot] = rank[t];
}

template <int N, class K>
__device__ inline void bitonic_regs(K (&a)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const K x = a[i], y = a[l];
          const bool sw = (x > y) == ((i & k) == 0);  // ascending where bit k of i is clear
          a[i] = sw ? y : x;
          a[l] = sw ? x : y;
        }
      }
}

// One thread per segment of up to N entries: sort (order key, slot content) in registers.
template <int N, bool EXACT>
__device__ inline void sort_segment_regs(uint32_t s, uint32_t len, uint32_t* perm, uint32_t* skey) {
  if (EXACT) {
    unsigned long long a[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
      a[i] = (uint32_t)i < len ? ((unsigned long long)skey[s + i] << 32) | perm[s + i] : ~0ull;
    bitonic_regs<N>(a);
#pragma unroll
    for (int i = 0; i < N; ++i)
      if ((uint32_t)i < len) {
        perm[s + i] = (uint32_t)a[i];
        skey[s + i] = (uint32_t)(a[i] >> 32);
      }
  } else {
    uint32_t a[N];
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = (uint32_t)i < len ? perm[s + i] : 0xffffffffu;
    bitonic_regs<N>(a);
#pragma unroll
    for (int i = 0; i < N; ++i)
      if ((uint32_t)i < len) perm[s + i] = a[i];
  }
}

template <bool EXACT>
__global__ void k_cell_order(const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ cend, int64_t nc,
                             uint32_t* __restrict__ perm, uint32_t* __restrict__ skey, uint32_t* __restrict__ big,
                             uint32_t* __restrict__ nbig) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t s = cstart[c], len = cend[c] - s;
  if (len <= 1) return;
  if (len > kRegSort) {
    big[atomicAdd(nbig, 1u)] = (uint32_t)c;
    return;
  }
  if (len <= 8)
    sort_segment_regs<8, EXACT>(s, len, perm, skey);
  else
    sort_segment_regs<16, EXACT>(s, len, perm, skey);
}

// Bitonic sort of `len` (<= kRun) 64-bit keys in LDS, padded to a power of two.
__device__ inline void lds_bitonic(unsigned long long* key, uint32_t len) {
  uint32_t P = 2;
  while (P < len) P <<= 1;
  for (uint32_t i = len + threadIdx.x; i < P; i += blockDim.x) key[i] = ~0ull;
  __syncthreads();
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const unsigned long long a = key[i], b = key[l];
          if ((a > b) == ((i & k) == 0)) {
            key[i] = b;
            key[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Sorts `len` (key, value) entries at kv/vv (+ off) in place: runs of kRun in LDS, then
// merge passes between (kv, vv) and (tk, tv) at the same offsets; an entry's place in a
// merged run is its index in its own run plus the number of smaller keys in the partner
// run (keys are unique).  An odd number of passes ends with a copy back.  kv == vv
// means the value is the key.
__device__ void block_sort_segment(uint32_t off, uint32_t len, uint32_t* kv, uint32_t* vv, uint32_t* tk,
                                   uint32_t* tv, unsigned long long* key) {
  const bool same = kv == vv;
  for (uint32_t r0 = 0; r0 < len; r0 += kRun) {
    const uint32_t rl = len - r0 < kRun ? len - r0 : kRun;
    for (uint32_t i = threadIdx.x; i < rl; i += blockDim.x)
      key[i] = ((unsigned long long)kv[off + r0 + i] << 32) | vv[off + r0 + i];
    lds_bitonic(key, rl);
    for (uint32_t i = threadIdx.x; i < rl; i += blockDim.x) {
      const unsigned long long x = key[i];
      kv[off + r0 + i] = (uint32_t)(x >> 32);
      if (!same) vv[off + r0 + i] = (uint32_t)x;
    }
    __syncthreads();
  }
  if (len <= kRun) return;
  uint32_t *sk = kv, *sv = vv, *dk = tk, *dv = same ? tk : tv;
  int passes = 0;
  for (uint32_t w = kRun; w < len; w <<= 1, ++passes) {
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
      const uint32_t x = sk[off + i];
      const uint32_t run = i / w, pair = (run & ~1u) * w, pb = (run ^ 1u) * w;
      uint32_t lo = pb, hi = pb < len ? (pb + w < len ? pb + w : len) : pb;
      while (lo < hi) {  // first entry of the partner run with a key above x
        const uint32_t m = (lo + hi) >> 1;
        if (sk[off + m] < x) lo = m + 1;
        else hi = m;
      }
      const uint32_t dst = pair + (i - run * w) + (lo - pb);
      dk[off + dst] = x;
      if (!same) dv[off + dst] = sv[off + i];
    }
    __syncthreads();
    uint32_t* t = sk;
    sk = dk, dk = t;
    t = sv, sv = dv, dv = t;
  }
  if (passes & 1) {
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) {
      kv[off + i] = sk[off + i];
      if (!same) vv[off + i] = sv[off + i];
    }
    __syncthreads();
  }
}

// The segments longer than a thread sorts: one workgroup each.
__global__ void __launch_bounds__(kRunThreads) k_cell_order_big(const uint32_t* __restrict__ big,
                                                                const uint32_t* __restrict__ nbig,
                                                                const uint32_t* __restrict__ cstart,
                                                                const uint32_t* __restrict__ cend, uint32_t* kv,
                                                                uint32_t* vv, uint32_t* tk, uint32_t* tv) {
  __shared__ unsigned long long key[kRun];
  const uint32_t nb = *nbig;
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const uint32_t c = big[b];
    const uint32_t s = cstart[c];
    block_sort_segment(s, cend[c] - s, kv, vv, tk, tv, key);
  }
}

// Flat sort: runs of kRun entries, one workgroup each.
__global__ void __launch_bounds__(kRunThreads) k_flat_runs(uint32_t* kv, uint32_t* vv, int64_t n) {
  __shared__ unsigned long long key[kRun];
  const int64_t r0 = (int64_t)blockIdx.x * kRun;
  const uint32_t rl = (uint32_t)(n - r0 < (int64_t)kRun ? n - r0 : (int64_t)kRun);
  const bool same = kv == vv;
  for (uint32_t i = threadIdx.x; i < rl; i += blockDim.x)
    key[i] = ((unsigned long long)kv[r0 + i] << 32) | vv[r0 + i];
  lds_bitonic(key, rl);
  for (uint32_t i = threadIdx.x; i < rl; i += blockDim.x) {
    const unsigned long long x = key[i];
    kv[r0 + i] = (uint32_t)(x >> 32);
    if (!same) vv[r0 + i] = (uint32_t)x;
  }
}

// Flat sort: one merge pass, sorted runs of width w into runs of 2w.
__global__ void k_flat_merge(const uint32_t* __restrict__ sk, const uint32_t* __restrict__ sv,
                             uint32_t* __restrict__ dk, uint32_t* __restrict__ dv, int64_t n, int64_t w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t x = sk[i];
  const int64_t run = i / w, pair = (run & ~(int64_t)1) * w, pb = (run ^ 1) * w;
  int64_t lo = pb, hi = pb < n ? (pb + w < n ? pb + w : n) : pb;
  while (lo < hi) {
    const int64_t m = (lo + hi) >> 1;
    if (sk[m] < x) lo = m + 1;
    else hi = m;
  }
  const int64_t dst = pair + (i - run * w) + (lo - pb);
  dk[dst] = x;
  if (dv) dv[dst] = sv[i];
}

hipError_t flat_sort(uint32_t* kv, uint32_t* vv, uint32_t* tk, uint32_t* tv, int64_t n, hipStream_t st) {
  if (n <= 1) return hipSuccess;
  const bool same = kv == vv;
  k_flat_runs<<<blocks_for(n, kRun), kRunThreads, 0, st>>>(kv, vv, n);
  uint32_t *sk = kv, *sv = vv, *dk = tk, *dv = tv;
  int passes = 0;
  for (int64_t w = kRun; w < n; w <<= 1, ++passes) {
    k_flat_merge<<<blocks_for(n, 256), 256, 0, st>>>(sk, same ? nullptr : sv, dk, same ? nullptr : dv, n, w);
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  if (passes & 1) {
    SDR_RET(hipMemcpyAsync(kv, sk, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    if (!same) SDR_RET(hipMemcpyAsync(vv, sv, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  }
  return hipGetLastError();
}

}  // namespace

size_t scan_u32_temp(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile + 1) * 4; }

hipError_t exclusive_sum_u32(void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out, int64_t n,
                             hipStream_t st) {
  return scan_launch<Ident, false>(temp, temp_bytes, in, Ident{}, out, nullptr, n, st);
}

hipError_t exclusive_sum_over_cap(void* temp, size_t temp_bytes, const uint32_t* count, uint32_t cap, bool sizes,
                                  uint32_t* out, int64_t n, hipStream_t st) {
  return scan_launch<OverCap, false>(temp, temp_bytes, count, OverCap{cap, sizes ? 1 : 0}, out, nullptr, n, st);
}

hipError_t sort_keys_u32(const uint32_t* kin, uint32_t* kout, uint32_t* scratch_k, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (kout != kin) SDR_RET(hipMemcpyAsync(kout, kin, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  return flat_sort(kout, kout, scratch_k, scratch_k, n, st);
}

hipError_t sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout,
                          uint32_t* scratch_k, uint32_t* scratch_v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (kout != kin) SDR_RET(hipMemcpyAsync(kout, kin, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  if (vout != vin) SDR_RET(hipMemcpyAsync(vout, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  return flat_sort(kout, vout, scratch_k, scratch_v, n, st);
}

template <class T>
void cell_sort(const Parts<T>& p, int64_t n, GridD<T> g, bool exact, const CellSortBufs& b, void* temp,
               size_t temp_bytes, hipStream_t st) {
  const int64_t nc = (int64_t)g.nx * g.ny * g.nz;
  fill_u32(b.cend, nc, 0u, st);
  fill_u32(b.cursor, nc, 0u, st);
  fill_u32(b.ctr, 1, 0u, st);
  if (n > 0) k_cell_count<T><<<blocks_for(n, 256), 256, 0, st>>>(p, n, g, b.key, b.cend);
  // cend holds the counts: cstart = their exclusive sums, cend = the inclusive ones
  HIPCHK((scan_launch<Ident, true>(temp, temp_bytes, b.cend, Ident{}, b.cstart, b.cend, nc, st)));
  if (n <= 0) return;
  k_cell_scatter<<<blocks_for(n, 256), 256, 0, st>>>(b.key, n, b.cstart, b.cursor, b.perm,
                                                     exact ? p.rank : nullptr, exact ? b.skey : nullptr);
  if (exact) {
    k_cell_order<true><<<blocks_for(nc, 256), 256, 0, st>>>(b.cstart, b.cend, nc, b.perm, b.skey, b.big, b.ctr);
    k_cell_order_big<<<1024, kRunThreads, 0, st>>>(b.big, b.ctr, b.cstart, b.cend, b.skey, b.perm, b.tskey,
                                                   b.tperm);
  } else {
    k_cell_order<false><<<blocks_for(nc, 256), 256, 0, st>>>(b.cstart, b.cend, nc, b.perm, nullptr, b.big, b.ctr);
    k_cell_order_big<<<1024, kRunThreads, 0, st>>>(b.big, b.ctr, b.cstart, b.cend, b.perm, b.perm, b.tperm,
                                                   b.tperm);
  }
  HIPCHK(hipGetLastError());
}

template void cell_sort<float>(const Parts<float>&, int64_t, GridD<float>, bool, const CellSortBufs&, void*, size_t,
                               hipStream_t);
template void cell_sort<double>(const Parts<double>&, int64_t, GridD<double>, bool, const CellSortBufs&, void*,
                                size_t, hipStream_t);

}  // namespace sdr
