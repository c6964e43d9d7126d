// redist.hip — the order-dependent decision pass of redistribute (solver.rs:454-557)
// as deterministic reservation rounds over cell buckets.
//
// During a sweep a cell over the cap (a source) only loses particles and only down to
// the cap, a cell under it (a receiver) only gains and only up to the cap, and a cell at
// the cap never changes.  So a source's particle reads only its own count and its
// receiver neighbours' counts, and only the first count - cap particles of a source in
// reference order can move (once one finds no receiver with room, none of the later
// ones will: rooms only shrink).  Sources interact only through shared receivers.
//
// Layout of one sweep:
//   * sources are numbered by an exclusive scan over the cells (count > cap), and each
//     gets a bucket of count entries at an offset from a second scan (a counting sort:
//     no radix sort of the particles);
//   * every particle of a source cell lands in its bucket; the bucket's first
//     count - cap entries are then put in reference (rank) order by a partial
//     selection, one thread per source (one wave for large clumps);
//   * rounds: the head particle of every pending source reserves its open receivers
//     with (round, reference rank) by a 64-bit atomic max; a head whose choice depends
//     only on receivers it holds (or that are already full) sees exactly the counts the
//     sequential loop would and commits.  The lowest pending head always commits, so
//     every round makes progress.  Pending sources are appended to the next round's
//     list with one atomic per block, so a round costs in proportion to what is still
//     pending.
//   * a committing source goes on with the particles after its head, in reference
//     order, while each choice still depends only on held or full receivers and the
//     rank stays below the earliest head that lost a held receiver to it this round
//     (the reservation keeps that rank too): no particle of another cell comes between
//     them in the sequential loop and touches these receivers, so a clump whose
//     downhill neighbours are uncontested drains in one round instead of one round per
//     particle.
#include <hip/hip_runtime.h>
#include <math.h>

#include "kernels.hpp"

namespace sdr {
namespace {

constexpr int kRdBlock = 256;
constexpr uint32_t kBig = 64;         // buckets above this size are sorted by a workgroup
constexpr uint32_t kSortRun = 8192;   // entries a workgroup sorts in LDS (64 KB of keys)
constexpr int kSortThreads = 1024;
constexpr int kRun = 8;        // bucket ranks a committing source loads at once

inline unsigned rd_grid(int64_t n, int block = kRdBlock) {
  int64_t b = (n + block - 1) / block;
  return (unsigned)(b < 1 ? 1 : b);
}

// Slot of this thread in a list appended to by every thread with `flag`; one atomic per
// block.  Every thread of the block must call it (it synchronises the block).
__device__ inline uint32_t block_append(uint32_t* ctr, bool flag) {
  __shared__ uint32_t wbase[kRdBlock / 64];
  __shared__ uint32_t bbase;
  const unsigned long long m = __ballot(flag);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wbase[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < kRdBlock / 64; ++w) {
      const uint32_t c = wbase[w];
      wbase[w] = tot;
      tot += c;
    }
    bbase = tot ? atomicAdd(ctr, tot) : 0u;
  }
  __syncthreads();
  const uint32_t r = bbase + wbase[wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  __syncthreads();
  return r;
}

// Sum of a per-thread count over the block, added to *ctr with one atomic.
__device__ inline void block_add64(unsigned long long* ctr, uint32_t v) {
  __shared__ uint32_t part[kRdBlock / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kRdBlock / 64; ++w) t += part[w];
    if (t) atomicAdd(ctr, t);
  }
}

__device__ inline int64_t rd_choose(int64_t c, const uint32_t* count, const uint8_t* mask, int nx,
                                    int ny, int nz, uint32_t cap, const RedistDirs& dn) {
  // solver.rs:486-544: visit -x,+x,-y,+y,-z,+z; most downhill, then most room, then first
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sz);
  const int64_t nb[6] = {c - 1, c + 1, c - sy, c + sy, c - sz, c + sz};
  const bool in[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
  bool have = false;
  int64_t best = -1;
  double best_down = -INFINITY;
  uint32_t best_room = 0;
  for (int q = 0; q < 6; ++q) {
    if (!in[q]) continue;
    const int64_t ncl = nb[q];
    if (!mask[ncl] || count[ncl] >= cap) continue;
    const uint32_t room = cap - count[ncl];
    if (!have || dn.d[q] > best_down || (dn.d[q] == best_down && room > best_room)) {
      have = true;
      best = ncl;
      best_down = dn.d[q];
      best_room = room;
    }
  }
  return best;
}

// ctr: [0] sources, [1] bucket entries, [2] pending A, [3] pending B, [4] big buckets,
// [6..7] moves (u64)
__global__ void k_rd_totals(const uint32_t* __restrict__ count, const uint32_t* __restrict__ sidx,
                            const uint32_t* __restrict__ boff, int64_t nc, uint32_t cap, uint32_t* ctr) {
  const uint32_t last = count[nc - 1];
  const uint32_t ns = sidx[nc - 1] + (last > cap ? 1u : 0u);
  ctr[0] = ns;
  ctr[1] = boff[nc - 1] + (last > cap ? last : 0u);
  ctr[2] = ns;
  ctr[3] = 0;
  ctr[4] = 0;
  ctr[5] = 0;
  *(unsigned long long*)(ctr + 6) = 0ull;
}

__global__ void k_rd_src_init(const uint32_t* __restrict__ count, const uint32_t* __restrict__ sidx,
                              const uint32_t* __restrict__ boff, int64_t nc, uint32_t cap,
                              uint32_t* src_cell, uint32_t* src_off, uint32_t* src_head,
                              uint32_t* src_lim, uint32_t* src_fill, uint32_t* pend, uint32_t* ctr5) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t cnt = count[c];
  if (cnt <= cap) return;
  const uint32_t s = sidx[c], off = boff[c];
  src_cell[s] = (uint32_t)c;
  src_off[s] = off;
  src_head[s] = off;
  src_lim[s] = off + (cnt - cap);
  src_fill[s] = 0;
  pend[s] = s;
  if (cnt > kSortRun) atomicMax(&ctr5[0], cnt);  // buckets that need merge passes (rare)
}

// Every particle of a source cell takes a slot of its cell's bucket.  Particles are
// close to cell order, so each run of equal cells in a wave takes its slots with one
// atomic; the order inside a bucket is fixed afterwards by reference rank.
__global__ void k_rd_bucket(const int32_t* __restrict__ cellkey, const uint32_t* __restrict__ count,
                            int64_t n, uint32_t cap, const uint32_t* __restrict__ sidx,
                            const uint32_t* __restrict__ src_off, uint32_t* src_fill,
                            const uint32_t* __restrict__ rank, uint32_t* brank, uint32_t* bslot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int key = -1;
  if (t < n) {
    const int c = cellkey[t];
    if (c >= 0 && count[c] > cap) key = c;
  }
  const int prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
  const int start = 63 - __clzll((long long)upto);
  uint32_t base = 0;
  uint32_t s = 0;
  if (key >= 0) s = sidx[key];
  if (head && key >= 0) {
    const unsigned long long after = lane == 63 ? 0ull : heads & ~((2ull << lane) - 1ull);
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    base = atomicAdd(&src_fill[s], (uint32_t)(end - lane));
  }
  base = __shfl(base, start);
  if (key < 0) return;
  const uint32_t pos = src_off[s] + base + (uint32_t)(lane - start);
  brank[pos] = rank[t];
  bslot[pos] = (uint32_t)t;
}

// The first count - cap entries of each bucket in reference order (partial selection;
// the order of the rest does not matter).  Buckets over kBig are sorted in full
// afterwards.
__global__ void k_rd_select(int64_t nsrc, const uint32_t* __restrict__ src_cell,
                            const uint32_t* __restrict__ src_off, const uint32_t* __restrict__ count,
                            uint32_t cap, uint32_t* brank, uint32_t* bslot, uint32_t* big,
                            uint32_t* nbig) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsrc) return;
  const uint32_t cnt = count[src_cell[s]];
  if (cnt > kBig) {
    big[atomicAdd(nbig, 1u)] = (uint32_t)s;
    return;
  }
  const uint32_t off = src_off[s], m = cnt - cap;
  for (uint32_t p = 0; p < m; ++p) {
    uint32_t bi = off + p, br = brank[bi];
    for (uint32_t q = off + p + 1; q < off + cnt; ++q) {
      const uint32_t r = brank[q];
      if (r < br) br = r, bi = q;
    }
    if (bi != off + p) {
      const uint32_t r0 = brank[off + p], s0 = bslot[off + p];
      brank[off + p] = br;
      bslot[off + p] = bslot[bi];
      brank[bi] = r0;
      bslot[bi] = s0;
    }
  }
}

// Large buckets are sorted by rank in full (the selection needs only the first
// count - cap entries, which in a clump is nearly all of them).  Ranks are unique, so the
// sorted order is fixed whatever the network.
//   * first, runs of up to kSortRun entries of every large bucket are sorted in LDS: a
//     bitonic network over (rank << 32 | slot), padded to a power of two;
//   * buckets longer than a run are then merged pass by pass, runs of width w into runs
//     of 2w, between the bucket arrays and a scratch copy: an entry's place in the merged
//     run is its index in its own run plus the number of smaller ranks in the partner run
//     (a binary search).  An odd number of passes ends with a copy back.
__global__ void __launch_bounds__(kSortThreads) k_rd_sort_runs(const uint32_t* __restrict__ nbig,
                                                               const uint32_t* __restrict__ big,
                                                               const uint32_t* __restrict__ src_cell,
                                                               const uint32_t* __restrict__ src_off,
                                                               const uint32_t* __restrict__ count,
                                                               uint32_t* brank, uint32_t* bslot) {
  __shared__ unsigned long long key[kSortRun];
  const uint32_t nb = *nbig;
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const uint32_t s = big[b], off = src_off[s], cnt = count[src_cell[s]];
    for (uint32_t r0 = 0; r0 < cnt; r0 += kSortRun) {
      const uint32_t len = cnt - r0 < kSortRun ? cnt - r0 : kSortRun;
      uint32_t P = 64;
      while (P < len) P <<= 1;
      for (uint32_t i = threadIdx.x; i < P; i += kSortThreads)
        key[i] = i < len ? ((unsigned long long)brank[off + r0 + i] << 32) | bslot[off + r0 + i] : ~0ull;
      __syncthreads();
      for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
          for (uint32_t i = threadIdx.x; i < P; i += kSortThreads) {
            const uint32_t l = i ^ j;
            if (l > i) {
              const unsigned long long a = key[i], c = key[l];
              if ((a > c) == ((i & k) == 0)) {  // ascending where bit k of i is clear
                key[i] = c;
                key[l] = a;
              }
            }
          }
          __syncthreads();
        }
      }
      for (uint32_t i = threadIdx.x; i < len; i += kSortThreads) {
        const unsigned long long v = key[i];
        brank[off + r0 + i] = (uint32_t)(v >> 32);
        bslot[off + r0 + i] = (uint32_t)v;
      }
      __syncthreads();
    }
  }
}

// One merge pass over the buckets longer than kSortRun: sorted runs of width w become
// sorted runs of width 2w in the output arrays (w = 0: a plain copy).
__global__ void __launch_bounds__(kSortThreads) k_rd_merge_runs(
    const uint32_t* __restrict__ nbig, const uint32_t* __restrict__ big, const uint32_t* __restrict__ src_cell,
    const uint32_t* __restrict__ src_off, const uint32_t* __restrict__ count, const uint32_t* __restrict__ rin,
    const uint32_t* __restrict__ sin, uint32_t* __restrict__ rout, uint32_t* __restrict__ sout, uint32_t w) {
  const uint32_t nb = *nbig;
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const uint32_t s = big[b], off = src_off[s], cnt = count[src_cell[s]];
    if (cnt <= kSortRun) continue;
    for (uint32_t i = threadIdx.x; i < cnt; i += kSortThreads) {
      const uint32_t r = rin[off + i];
      uint32_t dst = i;
      if (w) {
        const uint32_t run = i / w, pair = (run & ~1u) * w;
        const uint32_t pb = (run ^ 1u) * w;  // the partner run
        uint32_t lo = pb, hi = pb < cnt ? (pb + w < cnt ? pb + w : cnt) : pb;
        while (lo < hi) {  // first entry of the partner run with a rank above r
          const uint32_t m = (lo + hi) >> 1;
          if (rin[off + m] < r) lo = m + 1;
          else hi = m;
        }
        dst = pair + (i - run * w) + (lo - pb);
      }
      rout[off + dst] = r;
      sout[off + dst] = sin[off + i];
    }
  }
}

// The first round's records: entry x of the pending list is source list[x] with its
// cell, head, limit and head rank.
__global__ void k_rd_rec_init(const uint32_t* __restrict__ list, const uint32_t* __restrict__ nlist,
                              const uint32_t* __restrict__ src_cell, const uint32_t* __restrict__ src_head,
                              const uint32_t* __restrict__ src_lim, const uint32_t* __restrict__ brank,
                              uint4* rec) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= *nlist) return;
  const uint32_t s = list[x], head = src_head[s], lim = src_lim[s];
  rec[x] = make_uint4(src_cell[s], head, lim, head < lim ? brank[head] : 0u);
}

// Reservation: a head reserves every open receiver it could still ever write (rooms only
// shrink, so later particles of its cell touch a subset); the set is kept for the commit.
// Entry x of the pending list comes as a record (cell, head, limit, head rank) left by
// the previous round's commit, read and kept in list order.
__global__ void __launch_bounds__(kRdBlock) k_rd_reserve2(
    const uint32_t* __restrict__ nlist, uint32_t* nnext, uint4* rec, uint8_t* rmx,
    const uint32_t* __restrict__ count, const uint8_t* __restrict__ mask, int nx, int ny, int nz,
    uint32_t cap, unsigned long long* resv, unsigned long long* second, uint32_t epoch) {
  const uint32_t nl = *nlist;
  if (blockIdx.x == 0 && threadIdx.x == 0) *nnext = 0;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < nl;
       x += (int64_t)gridDim.x * blockDim.x) {
    const uint4 e = rec[x];
    const uint32_t head = e.y, lim = e.z;
    if (head >= lim) {
      rmx[x] = 0;
      continue;
    }
    const int64_t c = e.x;
    const uint32_t hrank = e.w;
    const unsigned long long val = ((unsigned long long)epoch << 32) | (0xFFFFFFFFu - hrank);
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / sz);
    const int64_t nb[6] = {c - 1, c + 1, c - sy, c + sy, c - sz, c + sz};
    const bool in[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
    uint8_t rm = 0;
    for (int q = 0; q < 6; ++q) {
      if (!in[q] || !mask[nb[q]] || count[nb[q]] >= cap) continue;
      rm |= (uint8_t)(1u << q);
      // of this head and the one it displaced (or that kept the receiver), the later
      // loses; the earliest loser of each receiver is kept in `second`
      const unsigned long long old = atomicMax(resv + nb[q], val);
      const unsigned long long loser = old > val ? val : old;
      if ((uint32_t)(loser >> 32) == epoch) atomicMax(second + nb[q], loser);
    }
    // no receiver left: no later particle can move
    if (rm == 0) rec[x] = make_uint4((uint32_t)c, lim, lim, hrank);
    rmx[x] = rm;
  }
}

// Commit: a source moves its head, then the particles after it in reference order,
// while each choice is exact and no particle of another cell can come between:
//   * a receiver this source holds changes only by its own moves this round, so its
//     room is exact while the particle's rank stays below the earliest head that lost
//     one of the held receivers to it (kept by the reservation);
//   * a receiver held by another source is written this round only by particles that
//     come before this source's head (its holder's run stops at the earliest loser,
//     which is at most this head), and counts only grow: read full, it is full for
//     every particle of this source; read with room, its room is unknown;
// so a choice is exact when every open receiver at least as downhill as the chosen one
// is held or read full.  The head's own choice among the receivers of its best gravity
// class follows the same test.  Sources still pending go to the next round's list.
__global__ void __launch_bounds__(kRdBlock) k_rd_commit2(
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ nlist, uint32_t* next,
    uint32_t* nnext, const uint4* __restrict__ rec, uint4* rec_next, const uint8_t* __restrict__ rmx,
    const uint32_t* __restrict__ brank, uint32_t* count, int nx, int ny, uint32_t cap, RedistDirs dn,
    const unsigned long long* __restrict__ resv, const unsigned long long* __restrict__ second,
    uint32_t epoch, int32_t* dest, unsigned long long* moved) {
  const uint32_t nl = *nlist;
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  uint32_t nmoved = 0;
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < nl; b0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = b0 + threadIdx.x;
    bool pend = false;
    uint32_t s = 0;
    uint4 enext = make_uint4(0u, 0u, 0u, 0u);  // the entry's record for the next round
    if (x < nl) {
      s = list[x];
      const uint4 e = rec[x];  // cell, head, lim, head's rank
      const uint32_t head = e.y, lim = e.z;
      if (head < lim) {
        const int64_t c = e.x;
        const unsigned long long val = ((unsigned long long)epoch << 32) | (0xFFFFFFFFu - e.w);
        const int64_t nb[6] = {c - 1, c + 1, c - sy, c + sy, c - sz, c + sz};
        // the receivers open at the reservation (a closed one stays closed)
        const uint32_t rm = rmx[x];
        uint32_t held = 0;
        unsigned long long comp = 0;  // earliest competing head on a held receiver
        uint32_t room[6];
        for (int q = 0; q < 6; ++q) {
          room[q] = 0;
          if (!((rm >> q) & 1u)) continue;
          if (resv[nb[q]] == val) {
            held |= 1u << q;
            const unsigned long long s2 = second[nb[q]];
            if ((uint32_t)(s2 >> 32) == epoch && s2 > comp) comp = s2;
          }
          // rooms as the sequential loop sees them (solver.rs:486-544), exact for held
          // receivers and at most what is left for the others
          const uint32_t cn = count[nb[q]];
          room[q] = cn < cap ? cap - cn : 0u;
        }
        const uint64_t rank_lim = comp ? (uint64_t)(0xFFFFFFFFu - (uint32_t)comp) : (1ull << 32);
        uint32_t add[6] = {0, 0, 0, 0, 0, 0};
        uint32_t p = head, prank = e.w;  // the next particle and its rank
        bool exhausted = false, go = true;
        while (go) {
          // the ranks of the next kRun bucket entries, loaded together rather than one
          // dependent load per moved particle
          uint32_t rk[kRun];
#pragma unroll
          for (int t = 0; t < kRun; ++t) {
            const uint32_t e = p + 1 + (uint32_t)t;
            rk[t] = e < lim ? brank[e] : 0u;
          }
#pragma unroll
          for (int t = 0; t < kRun; ++t) {
            if (!go) continue;
            int bq = -1;
            double bd = -INFINITY;
            uint32_t br = 0;
            for (int q = 0; q < 6; ++q) {
              if (room[q] == 0) continue;
              if (bq < 0 || dn.d[q] > bd || (dn.d[q] == bd && room[q] > br)) bq = q, bd = dn.d[q], br = room[q];
            }
            if (bq < 0) {
              // every open receiver is full for good: no later particle can move
              exhausted = true;
              go = false;
              continue;
            }
            bool exact = true;
            for (int q = 0; q < 6; ++q)
              if (!((held >> q) & 1u) && room[q] != 0 && dn.d[q] >= bd) exact = false;
            if (!exact) {
              go = false;
              continue;
            }
            room[bq] -= 1;
            add[bq] += 1;
            dest[p] = (int32_t)nb[bq];
            ++p;
            // p is entry p0 + t + 1 now: rk[t] is its rank when it is below lim
            prank = rk[t];
            if (p >= lim || (uint64_t)prank >= rank_lim) go = false;
          }
        }
        const uint32_t nm = p - head;
        nmoved += nm;
        if (nm) {
          count[c] -= nm;
          for (int q = 0; q < 6; ++q)
            if (add[q]) count[nb[q]] += add[q];
        }
        pend = !exhausted && p < lim;
        enext = make_uint4((uint32_t)c, p, lim, prank);
      }
    }
    const uint32_t q = block_append(nnext, pend);
    if (pend) next[q] = s, rec_next[q] = enext;
  }
  block_add64(moved, nmoved);
}

}  // namespace

void launch_rd_prepare(const uint32_t* count, const uint32_t* sidx, const uint32_t* boff, int64_t nc,
                       uint32_t cap, uint32_t* src_cell, uint32_t* src_off, uint32_t* src_head,
                       uint32_t* src_lim, uint32_t* src_fill, uint32_t* pend, uint32_t* ctr,
                       hipStream_t st) {
  k_rd_totals<<<1, 1, 0, st>>>(count, sidx, boff, nc, cap, ctr);
  k_rd_src_init<<<rd_grid(nc), kRdBlock, 0, st>>>(count, sidx, boff, nc, cap, src_cell, src_off, src_head,
                                                  src_lim, src_fill, pend, ctr + 5);
}

void launch_rd_buckets(const int32_t* cellkey, const uint32_t* count, int64_t n, uint32_t cap,
                       const uint32_t* sidx, int64_t nsrc, const uint32_t* src_cell,
                       const uint32_t* src_off, uint32_t* src_fill, const uint32_t* rank,
                       uint32_t* brank, uint32_t* bslot, uint32_t* trank, uint32_t* tslot,
                       uint32_t* big, uint32_t* ctr, uint32_t max_count, hipStream_t st) {
  if (n > 0)
    k_rd_bucket<<<rd_grid(n), kRdBlock, 0, st>>>(cellkey, count, n, cap, sidx, src_off, src_fill, rank,
                                                 brank, bslot);
  if (nsrc <= 0) return;
  k_rd_select<<<rd_grid(nsrc), kRdBlock, 0, st>>>(nsrc, src_cell, src_off, count, cap, brank, bslot, big,
                                                  ctr + 4);
  k_rd_sort_runs<<<512, kSortThreads, 0, st>>>(ctr + 4, big, src_cell, src_off, count, brank, bslot);
  int passes = 0;
  for (uint64_t w = kSortRun; w < max_count; w <<= 1, ++passes) {
    const bool fwd = (passes & 1) == 0;
    k_rd_merge_runs<<<512, kSortThreads, 0, st>>>(ctr + 4, big, src_cell, src_off, count, fwd ? brank : trank,
                                                  fwd ? bslot : tslot, fwd ? trank : brank, fwd ? tslot : bslot,
                                                  (uint32_t)w);
  }
  if (passes & 1)
    k_rd_merge_runs<<<512, kSortThreads, 0, st>>>(ctr + 4, big, src_cell, src_off, count, trank, tslot, brank,
                                                  bslot, 0u);
}

void launch_rd_round2(int64_t max_pending, int parity, bool first, uint32_t* pend_a, uint32_t* pend_b,
                      const uint32_t* src_cell, const uint32_t* src_head, const uint32_t* src_lim,
                      uint32_t* rec_a, uint32_t* rec_b, uint8_t* rmx, const uint32_t* brank, uint32_t* count,
                      const uint8_t* mask,
                      int nx, int ny, int nz, uint32_t cap, RedistDirs dn, unsigned long long* resv,
                      unsigned long long* second, uint32_t epoch, int32_t* dest, uint32_t* ctr,
                      hipStream_t st) {
  if (max_pending <= 0) return;
  const uint32_t* in = parity ? pend_b : pend_a;
  uint32_t* out = parity ? pend_a : pend_b;
  uint32_t* nin = ctr + 2 + parity;
  uint32_t* nout = ctr + 3 - parity;
  int64_t blocks = (max_pending + kRdBlock - 1) / kRdBlock;
  if (blocks > 4096) blocks = 4096;
  // the records pair with the lists: rec_a with pend_a, rec_b with pend_b
  uint4* rin = reinterpret_cast<uint4*>(parity ? rec_b : rec_a);
  uint4* rout = reinterpret_cast<uint4*>(parity ? rec_a : rec_b);
  if (first)
    k_rd_rec_init<<<rd_grid(max_pending), kRdBlock, 0, st>>>(in, nin, src_cell, src_head, src_lim, brank, rin);
  k_rd_reserve2<<<(unsigned)blocks, kRdBlock, 0, st>>>(nin, nout, rin, rmx, count, mask, nx, ny, nz, cap, resv,
                                                       second, epoch);
  k_rd_commit2<<<(unsigned)blocks, kRdBlock, 0, st>>>(in, nin, out, nout, rin, rout, rmx, brank, count, nx, ny,
                                                      cap, dn, resv, second, epoch, dest,
                                                      (unsigned long long*)(ctr + 6));
}

}  // namespace sdr
