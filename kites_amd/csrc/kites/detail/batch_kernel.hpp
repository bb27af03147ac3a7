// batch_kernel.hpp -- the K-wide multi-source iteration kernel.
//
// Up to 64 independent sources run the same min-semilattice fixpoint at
// once.  Values are rows of 64 lanes (V[v * 64 + k] = value of vertex v for
// source k, 256 B per row for 4-byte values), so one wave handles one
// (edge, all sources) pair per instruction: lane k owns source k.  The
// frontier is one bit per vertex (A: some source's value of the vertex
// moved in the previous iteration); it is 512 KB at 2^22 vertices and
// stays in L2, so testing it per edge costs no HBM traffic.
//
// One launch = one iteration, pull over the in-edge CSR (C, E in
// destination order), atomic-free for vertices owned by one wave:
//   * the in-edge array is cut into fixed work items of IPL 64-edge chunks
//     (IPL = items per lane); wave w takes items w, w + waves, ...; each
//     chunk's first vertex position in the non-empty vertex list is
//     precomputed (istart, one entry per 64 edges);
//   * all loads of an item are issued first: lane j of chunk i loads C[e]
//     and E[e] (coalesced), then its source's frontier bit, and the chunk's
//     window of vertex ends; the edge's destination comes from a merge of
//     the chunk with the vertex ends through a per-wave LDS head array and
//     a wave max-scan;
//   * for every edge whose source is on the frontier, in destination order,
//     the wave loads the source's whole row (one coalesced 256-B access,
//     `UNROLL` rows in flight), applies nbrComp per lane and keeps the
//     running min per lane.  Relaxing a lane whose value did not move is a
//     harmless extra relaxation and costs no extra sector: a row is read as
//     a whole either way once several of its sources are active;
//   * on a destination change the row is flushed: a vertex whose edges all
//     lie inside the item is owned by this wave (plain compare and store);
//     a vertex cut by an item boundary is flushed with per-lane atomicMin.
//     A row that moved sets its frontier bit (atomicOr on the 64-bit word).
// The frontier ring has three slots: iteration i reads slot i%3, writes
// slot (i+1)%3 and clears slot (i+2)%3.
//
// Per-source masks (`M`, opt-in): besides the one-bit summary, a 64-bit
// word per vertex says which sources' values moved; lane k then relaxes an
// active edge only if bit k of its source's word is set, and a flushed row
// stores the ballot of its moved lanes.  The row is still read whole (any
// set bit needs it), so this design adds 8 B per active edge and per moved
// row and saves only arithmetic; it is kept as an option for that A/B.  The last workgroup of a launch
// (hierarchical ticket, as in the single-source kernel) ends the run when
// no vertex moved.
#ifndef KITES_DETAIL_BATCH_KERNEL_HPP_
#define KITES_DETAIL_BATCH_KERNEL_HPP_

#include <hip/hip_runtime.h>

#include <limits>
#include <type_traits>

#include "step_kernel.hpp"
#include "wave.hpp"

namespace kites {
namespace detail {

constexpr unsigned kBatchLanes = 64;

struct batch_stats {
  unsigned updated;            // vertices whose row moved (a cut vertex may count twice)
  unsigned flushes;            // rows flushed
  unsigned long long active_edges;   // in-edges whose source mask was non-zero
  unsigned long long gathered;       // source values loaded (set mask bits over those edges)
  unsigned long long moved;          // (vertex, source) values that moved
};

struct batch_ctl {
  unsigned iter;
  unsigned done;
  unsigned max_iter;
  unsigned overflow;
  unsigned push;               // bit-parallel BFS: the coming level runs top-down
  ticket_board board;          // completion-ticket counters (not copied to the host)
};

template <class vT, class eT, class FN>
struct batch_params {
  unsigned n;
  unsigned m;
  const unsigned* C;
  eT* E;
  const unsigned* nzi;
  const unsigned* nzR;
  unsigned nnz;
  const unsigned* istart;
  unsigned nitems;
  unsigned epi;
  vT* V;
  unsigned long long* A;       // 3 slots of aw words: frontier bit per vertex
  unsigned aw;
  unsigned long long* M;       // per-source masks: 3 slots of n words, or null
  batch_ctl* ctl;
  batch_stats* stats;
  bool collect;
  FN fn;
  batch_params(FN f) : fn(f) {}
};

__device__ __forceinline__ unsigned wave_incl_max(unsigned x) {
  unsigned y;
  y = dpp<0x111>(x); x = x > y ? x : y;
  y = dpp<0x112>(x); x = x > y ? x : y;
  y = dpp<0x114>(x); x = x > y ? x : y;
  y = dpp<0x118>(x); x = x > y ? x : y;
  y = dpp<0x142, 0xa>(x); x = x > y ? x : y;
  y = dpp<0x143, 0xc>(x); x = x > y ? x : y;
  return x;
}

__device__ __forceinline__ unsigned long long readlane64(unsigned long long x, unsigned lane) {
  unsigned const lo = __builtin_amdgcn_readlane(static_cast<unsigned>(x), lane);
  unsigned const hi = __builtin_amdgcn_readlane(static_cast<unsigned>(x >> 32), lane);
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Item start positions: istart[k] = position q of the non-empty vertex
// list whose edge range holds edge k * epi.
static __global__ void batch_istart_kernel(const unsigned* __restrict__ nzR, unsigned nnz, unsigned epi,
                                    unsigned nitems, unsigned* __restrict__ istart) {
  unsigned const k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nitems) return;
  unsigned const e = k * epi;
  unsigned lo = 0, hi = nnz;   // largest q with nzR[q] <= e
  while (hi - lo > 1) {
    unsigned const mid = (lo + hi) >> 1;
    if (nzR[mid] <= e) lo = mid; else hi = mid;
  }
  istart[k] = lo;
}

template <class vT>
__global__ void batch_fill_kernel(vT* __restrict__ V, unsigned long long rows, vT fill,
                                  unsigned long long* __restrict__ F, unsigned long long fwords) {
  unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long const stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
  for (unsigned long long j = i; j < rows * kBatchLanes; j += stride) V[j] = fill;
  for (unsigned long long j = i; j < fwords; j += stride) F[j] = 0ull;
}

template <class vT>
__global__ void batch_seed_kernel(vT* __restrict__ V, unsigned long long* __restrict__ A,
                                  const unsigned* __restrict__ src, unsigned K, vT value,
                                  batch_ctl* ctl, unsigned max_iter,
                                  unsigned long long* __restrict__ M) {
  unsigned const k = threadIdx.x;
  if (k < K) {
    unsigned const s = src[k];
    V[static_cast<unsigned long long>(s) * kBatchLanes + k] = value;
    atomicOr(A + (s >> 6), 1ull << (s & 63));
    if (M) atomicOr(M + s, 1ull << k);
  }
  if (k == 0) {
    ctl->iter = 0;
    ctl->done = K == 0;
    ctl->max_iter = max_iter;
    ctl->overflow = 0;
  }
  clear_ticket_board(ctl->board.w, k, blockDim.x);
}

// Applies a destination's row of minima.  `cur` is the row's value read
// when the destination was first seen: exact for a row this wave owns (no
// other wave writes it), an upper bound for a cut row.
template <class P, class vT>
__device__ __forceinline__ void batch_flush(P& p, unsigned long long* A_out, unsigned v, bool cut,
                                            vT acc, vT cur, unsigned lane, unsigned& upd_cnt,
                                            unsigned long long& moved,
                                            unsigned long long* M_out) {
  vT* row = p.V + static_cast<unsigned long long>(v) * kBatchLanes;
  bool up = false;
  if (!cut) {
    up = acc < cur;
    if (up) row[lane] = acc;
  } else if (acc < cur) {
    vT const old = atomic_min_any(row + lane, acc);
    up = acc < old;
  }
  unsigned long long const b = ballot64(up);
  if (b) {
    if (lane == 0) {
      atomicOr(A_out + (v >> 6), 1ull << (v & 63));
      if (M_out) {
        if (cut) atomicOr(M_out + v, b);
        else M_out[v] = b;
      }
    }
    ++upd_cnt;
    moved += __popcll(b);
  }
}

// PS: the per-source-mask variant (its own instantiation: the mask
// registers would cost the default kernel occupancy).
template <int BLOCK, int UNROLL, int IPL, bool PS, class P>
__global__ __launch_bounds__(BLOCK) void batch_step_kernel(P p) {
  using vT = typename std::remove_reference<decltype(*p.V)>::type;
  constexpr unsigned kW = BLOCK / 64;
  __shared__ unsigned heads[kW][64];
  // per wave and chunk of the item: source, destination (bit 31: the
  // destination is cut by an item boundary)
  __shared__ unsigned eu[kW][IPL][64];
  __shared__ unsigned ed[kW][IPL][64];
  // the chunk's edge weights, loaded coalesced with the sources; nbrComp
  // reads its weight through a pointer into this array
  using eT = typename std::remove_reference<decltype(*p.E)>::type;
  __shared__ eT ew[kW][IPL][64];
  __shared__ vT slot[BLOCK];
  __shared__ unsigned is_last;
  if (p.ctl->done) return;
  unsigned const iter = p.ctl->iter;
  unsigned const n = p.n;
  const unsigned long long* A_in = p.A + static_cast<unsigned long long>(iter % 3) * p.aw;
  unsigned long long* A_out = p.A + static_cast<unsigned long long>((iter + 1) % 3) * p.aw;
  unsigned long long* A_clr = p.A + static_cast<unsigned long long>((iter + 2) % 3) * p.aw;
  unsigned const t = threadIdx.x, lane = lane_id(), wib = t >> 6;
  const unsigned long long* M_in = nullptr;
  unsigned long long* M_out = nullptr;
  if constexpr (PS) {
    // per-source masks: clear the slot iteration iter + 1 writes where it
    // was written (the vertices its summary marks), one wave per word
    unsigned long long const ms = static_cast<unsigned long long>(n);
    M_in = p.M + (iter % 3) * ms;
    M_out = p.M + ((iter + 1) % 3) * ms;
    unsigned long long* M_clr = p.M + ((iter + 2) % 3) * ms;
    // (the wave that reads a summary word also zeroes it)
    for (unsigned w = blockIdx.x * kW + wib; w < p.aw; w += gridDim.x * kW) {
      unsigned long long const word = A_clr[w];
      if ((word >> lane) & 1ull) M_clr[static_cast<unsigned long long>(w) * 64u + lane] = 0ull;
      if (word && lane == 0) A_clr[w] = 0ull;
    }
  } else {
    for (unsigned i = blockIdx.x * BLOCK + t; i < p.aw; i += gridDim.x * BLOCK) A_clr[i] = 0ull;
  }

  vT const vmax = std::numeric_limits<vT>::max();
  unsigned upd_cnt = 0, flushes = 0;
  unsigned long long act_edges = 0, gathered = 0, moved = 0;
  unsigned const nw = gridDim.x * kW;
  constexpr unsigned kEpi = 64u * IPL;
  for (unsigned k = blockIdx.x * kW + wib; k < p.nitems; k += nw) {
    unsigned const e0 = k * kEpi;
    unsigned const e1 = min(e0 + kEpi, p.m);
    // Phase 1: every load of the item's chunks is issued before any is used
    // (sources, weights, frontier words, the chunk's non-empty-vertex
    // window), so the item costs a few memory round trips instead of a few
    // per chunk.
    unsigned u[IPL], vpv[IPL], endv[IPL], startv[IPL], vidv[IPL];
    bool valid[IPL];
#pragma unroll
    for (int i = 0; i < IPL; ++i) {
      unsigned const b = e0 + 64u * i;
      unsigned const e = b + lane;
      valid[i] = e < e1;
      u[i] = valid[i] ? p.C[e] : 0u;
      ew[wib][i][lane] = valid[i] ? p.E[e] : eT{};
      unsigned const c = b < e1 ? (b >> 6) : (e0 >> 6);
      vpv[i] = p.istart[c];
      unsigned const pos = vpv[i] + lane;
      endv[i] = pos < p.nnz ? p.nzR[pos + 1] : 0xffffffffu;
      startv[i] = pos < p.nnz ? p.nzR[pos] : 0xffffffffu;
      vidv[i] = pos < p.nnz ? p.nzi[pos] : 0u;
    }
    bool on_src[IPL];
#pragma unroll
    for (int i = 0; i < IPL; ++i)
      on_src[i] = valid[i] && ((A_in[u[i] >> 6] >> (u[i] & 63)) & 1ull);
    // Phase 2: destinations (merge of the chunk with the vertex ends).
#pragma unroll
    for (int i = 0; i < IPL; ++i) {
      unsigned const b = e0 + 64u * i;
      heads[wib][lane] = 0;
      wave_lds_sync();
      if (endv[i] > b && endv[i] - b < 64u) atomicMax(&heads[wib][endv[i] - b], lane + 1);
      wave_lds_sync();
      unsigned const off = wave_incl_max(heads[wib][lane]);
      unsigned const dst = __shfl(vidv[i], static_cast<int>(off), 64);
      unsigned const cut = __shfl(static_cast<unsigned>(startv[i] < e0 || endv[i] > e1),
                                  static_cast<int>(off), 64);
      eu[wib][i][lane] = on_src[i] ? u[i] : 0xffffffffu;
      ed[wib][i][lane] = dst | (cut << 31);
    }
    wave_lds_sync();
    // Phase 3: the active edges in destination order, `UNROLL` neighbour
    // rows in flight; flush a row when the destination changes.
    unsigned cur = 0xffffffffu;
    bool cur_cut = false;
    vT acc = vmax, oldv = vmax;
#pragma unroll 1
    for (int i = 0; i < IPL; ++i) {
      // chunk i's active edges (read back from LDS: a runtime index into the
      // register arrays would spill them)
      unsigned long long act = ballot64(eu[wib][i][lane] != 0xffffffffu);
      if (p.collect) act_edges += __popcll(act);
      unsigned const b = e0 + 64u * i;
      while (act) {
        unsigned js[UNROLL];
        unsigned cnt = 0;
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
          if (act) {
            js[q] = __builtin_ctzll(act);
            act &= act - 1;
            ++cnt;
          } else {
            js[q] = js[q > 0 ? q - 1 : 0];
          }
        }
        vT x[UNROLL];
        bool use[UNROLL];
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
          unsigned const uj = eu[wib][i][js[q]];
          x[q] = p.V[static_cast<unsigned long long>(uj) * kBatchLanes + lane];
          use[q] = true;
          if constexpr (PS) use[q] = (M_in[uj] >> lane) & 1ull;
        }
        if (p.collect && lane == 0) gathered += static_cast<unsigned long long>(cnt) * kBatchLanes;
#pragma unroll
        for (int q = 0; q < UNROLL; ++q) {
          if (static_cast<unsigned>(q) < cnt) {
            unsigned const j = js[q];
            unsigned const dc = ed[wib][i][j];
            unsigned const d = dc & 0x7fffffffu;
            if (d != cur) {
              if (cur != 0xffffffffu) {
                batch_flush(p, A_out, cur, cur_cut, acc, oldv, lane, upd_cnt, moved, M_out);
                ++flushes;
              }
              cur = d;
              cur_cut = (dc >> 31) != 0u;
              acc = vmax;
              // the row's current value, read now and used at its flush
              oldv = p.V[static_cast<unsigned long long>(d) * kBatchLanes + lane];
            }
            if (use[q]) {
              p.fn(vs(slot[t]), x[q], &ew[wib][i][j]);
              vT const c = vs(slot[t]);
              acc = c < acc ? c : acc;
            }
          }
        }
      }
    }
    if (cur != 0xffffffffu) {
      batch_flush(p, A_out, cur, cur_cut, acc, oldv, lane, upd_cnt, moved, M_out);
      ++flushes;
    }
    wave_lds_sync();
  }

  // Per-wave counters, then the launch's ticket.
  if (lane == 0) {
    batch_stats* st = p.stats + iter;
    if (upd_cnt) keep_live(atomicAdd(&st->updated, upd_cnt));
    if (p.collect) {
      if (flushes) keep_live(atomicAdd(&st->flushes, flushes));
      if (act_edges) keep_live(atomicAdd(&st->active_edges, act_edges));
      if (moved) keep_live(atomicAdd(&st->moved, moved));
    }
  }
  if (p.collect) {
    unsigned long long const g = wave_sum64(gathered);
    if (lane == 0 && g) keep_live(atomicAdd(&p.stats[iter].gathered, g));
  }
  __syncthreads();
  if (t == 0) {
    __builtin_amdgcn_s_waitcnt(0);
    is_last = take_ticket(p.ctl->board.w);
  }
  __syncthreads();
  if (!is_last || t != 0) return;
  unsigned const upd = __hip_atomic_load(&p.stats[iter].updated, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
  p.ctl->iter = iter + 1;
  if (upd == 0) p.ctl->done = 1;
  else if (iter + 1 >= p.ctl->max_iter) {
    p.ctl->done = 1;
    p.ctl->overflow = 1;
  }
}

// Order-sensitive 64-bit digest of the first K columns of n rows of u32
// values `stride` apart (K-wide rows: 64; bit-parallel levels: 256): the sum
// over entries of a mix of (position, value), into *out.  Lets an autotuner
// compare the results of two shapes without copying them to the host.
__device__ __forceinline__ unsigned long long digest_mix(unsigned long long x) {
  // a 64-bit finalizer (xorshift-multiply rounds)
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

static __global__ void rows_digest_kernel(const unsigned* __restrict__ L, unsigned n,
                                          unsigned stride, unsigned K, unsigned long long* out) {
  unsigned long long acc = 0ull;
  unsigned long long const total = static_cast<unsigned long long>(n) * stride;
  for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < total; i += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
    if (i % stride < K) acc += digest_mix((i << 32) ^ L[i]);
  }
  acc = wave_sum64(acc);
  if (lane_id() == 0 && acc) atomicAdd(out, acc);
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_BATCH_KERNEL_HPP_
