// persist_kernel.hpp -- small push frontiers in one workgroup, many
// iterations per launch.
//
// High-diameter graphs (grids, road-like meshes, paths) run thousands of
// iterations whose queues hold a few hundred or thousand vertices.  A launch
// of step_kernel per iteration then pays its dispatch, the completion ticket
// and the next decision -- 8 to 12 us on MI355X -- for very little work.
// persist_kernel runs such iterations in one workgroup of 1,024 threads that
// synchronizes with barriers only (there is no other workgroup, so nothing
// waits on one), until the queue is empty (the run is done), the schedule
// picks a variant that is not a push, the queue outgrows persist_items()
// entries or the run's out-edge limit (the next step_kernel launch takes
// over), or the iteration limit is reached.  One compute unit issues about
// one scattered access per cycle, so the loop pays off on queues of up to
// about 2,048 out-edges (the default limit; measured on grids and paths).
// The engine launches it first in every chunk it replays after a tiny push
// frontier; when the coming iteration does not fit, it returns at once.
//
// Iteration j uses the rings of step_kernel (queue slot j & 1, bitmap slot
// j % 3, the statistics entry j, the control block), so step_kernel and
// persist_kernel iterations follow each other in any order:
//   1. zero the words of bitmap slot (j + 2) % 3 that hold the queue
//      iteration j - 2 appended (queue slot (j + 1) & 1, read before this
//      iteration appends there): that slot was all zero before j - 2 wrote
//      it, and every bit j - 2 set belongs to a vertex it appended;
//      iteration j + 1 writes the slot;
//   2. stage the queue's out-edge offsets, the entries' values and their
//      first out-edges in LDS;
//   3. every thread takes edges e with e % 1024 = its index, kEPT of them
//      per pass, finds each edge's entry by binary search over the staged
//      offsets and issues the loads of a pass before it uses any; a
//      relaxation is a returning atomic min, and a destination it lowers
//      that has out-edges claims its bit in the next bitmap with a returning
//      atomic or (the bit decides the append, as in step_kernel); appended
//      vertices reserve queue slots and out-edge offsets with one 64-bit LDS
//      atomic per wave, so queue order and offset order agree;
//   4. close the next queue (sentinel offset), write the statistics and
//      decide the next variant with the schedule (choose_variant).
//
// Near-far runs (NF, unsigned values): a relaxation to a value at or above
// the threshold T claims the destination's far bit instead and appends it
// to the current far pile (its counter lives in a register of this launch:
// no other workgroup appends while the loop runs), and the least such value
// is kept.  When the queue runs dry with entries on the pile, the next
// iteration is a split (V_SPLIT, as in step_kernel): T becomes the least
// far value + delta, and every pile entry is dropped (its value fell below
// the previous threshold: it went through the queue), moved to the queue
// (below T) or kept on the other pile; the loop runs splits of up to
// 2 x the out-edge limit entries.
//
// The compute unit's L1 keeps lines across the iterations of one launch and
// atomics update only the L2, so every read of a location that atomics
// write (V, the bitmaps) is an agent-scope atomic load.  Plain stores of
// this workgroup (queues, offsets, piles, bitmap clears) pass through the
// same L1.
#ifndef KITES_DETAIL_PERSIST_KERNEL_HPP_
#define KITES_DETAIL_PERSIST_KERNEL_HPP_

#include <hip/hip_runtime.h>

#include <type_traits>

#include "step_kernel.hpp"
#include "wave.hpp"

namespace kites {
namespace detail {

constexpr unsigned kPersistBlock = 1024;
constexpr unsigned kEPT = 8;              // edges (or pile entries) per thread and pass
// Largest out-edge total of a queue it takes (step_core::persist_edges
// sets the run's limit, at most this).
constexpr unsigned kPersistEdges = 16384;
constexpr unsigned kPersistDefaultEdges = 2048;
// Queue entries staged in LDS (offset, value, first out-edge: 12 or 16 B).
template <class vT>
__host__ __device__ constexpr unsigned persist_items() {
  return sizeof(vT) <= 4 ? 4096u : 2048u;
}

template <class vT>
struct persist_smem {
  static constexpr unsigned kItems = persist_items<vT>();
  unsigned off[kItems + 1];   // out-edge offsets of the queue, sentinel last
  vT val[kItems];             // the entries' values
  unsigned base[kItems];      // the entries' first out-edge
  vT slot[kPersistBlock];     // nbrComp's output, one per thread
  unsigned long long qe;      // packed (count, out-edges) of the next queue
  unsigned long long vmin;    // near-far: least far value of this iteration
  unsigned maxdeg;
  unsigned far_add;           // near-far: entries appended to a pile
  unsigned go;                // 1: the next iteration runs here too
  unsigned next;              // the next iteration's variant
  float mlp[MLP_SCRATCH];     // the network's weights for an inference
};

// Whether iteration `iter` (variant `v`; a queue of `count` entries with
// `sdeg` out-edges, or a split of a pile of `far_n` entries) is one
// persist_kernel runs.
template <bool FAR, class vT>
__device__ __forceinline__ bool persist_fits(unsigned iter, unsigned v, unsigned count,
                                             unsigned long long sdeg, unsigned far_n,
                                             unsigned max_edges) {
  if (iter < 2) return false;
  if (v == V_SPLIT) return FAR && far_n <= 2u * max_edges;
  return v >= V_PUSH_T && v <= V_PUSH_LB && count <= persist_items<vT>() && sdeg <= max_edges;
}

// Largest i < count with off[i] <= e (off non-decreasing, off[0] = 0 <= e),
// by binary lifting over at most kItems entries.
template <unsigned kItems>
__device__ __forceinline__ unsigned persist_owner(const unsigned* off, unsigned count, unsigned e) {
  unsigned lo = 0;
#pragma unroll
  for (unsigned step = kItems / 2; step > 0; step >>= 1)
    if (lo + step < count && off[lo + step] <= e) lo += step;
  return lo;
}

template <class vT>
__device__ __forceinline__ vT load_agent(const vT* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Appends the lanes' vertices with `want` set (out-degree `d`) to the next
// queue: one 64-bit LDS atomic per wave reserves the slots and the out-edge
// offsets.  All 64 lanes active.
template <class vT>
__device__ __forceinline__ void persist_append(persist_smem<vT>& sm, bool want, unsigned v,
                                               unsigned d, unsigned* out_q, unsigned* out_qe,
                                               unsigned& wmax) {
  unsigned long long const m = ballot64(want);
  if (!m) return;
  unsigned const dd = want ? d : 0u;
  unsigned const incl = wave_incl_scan(dd);
  unsigned const dsum = readlane(incl, 63);
  unsigned long long old = 0;
  if (lane_id() == 0)
    old = atomicAdd(&sm.qe, qe_pack(static_cast<unsigned>(__builtin_popcountll(m)), dsum));
  unsigned const olo = readlane(static_cast<unsigned>(old), 0);
  unsigned const ohi = readlane(static_cast<unsigned>(old >> 32), 0);
  unsigned long long const o64 = (static_cast<unsigned long long>(ohi) << 32) | olo;
  if (want) {
    unsigned const pos = qe_count(o64) + rank_in(m);
    out_q[pos] = v;
    out_qe[pos] = qe_edges(o64) + incl - dd;
  }
  wmax = dd > wmax ? dd : wmax;
}

// Appends the lanes' vertices with `want` set to a far pile at `base` + a
// position reserved with one LDS atomic per wave.  All 64 lanes active.
template <class vT>
__device__ __forceinline__ void persist_far_append(persist_smem<vT>& sm, bool want, unsigned v,
                                                   unsigned* pile, unsigned base) {
  unsigned long long const m = ballot64(want);
  if (!m) return;
  unsigned old = 0;
  if (lane_id() == 0) old = atomicAdd(&sm.far_add, static_cast<unsigned>(__builtin_popcountll(m)));
  old = readlane(old, 0);
  if (want) pile[base + old + rank_in(m)] = v;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    unsigned long long const y = __shfl_xor(x, s, 64);
    x = y < x ? y : x;
  }
  return x;
}

template <bool NF, class P>
__global__ __launch_bounds__(kPersistBlock) void persist_kernel(P p) {
  using vT = typename std::remove_reference<decltype(*p.V)>::type;
  constexpr bool FAR = NF && std::is_unsigned<vT>::value;
  constexpr unsigned kItems = persist_items<vT>();
  __shared__ persist_smem<vT> sm;
  unsigned const t = threadIdx.x, lane = lane_id(), wib = t >> 6;
  if (p.ctl->done) return;
  if (p.sched.mode != SCHED_PLAIN && p.sched.mode != SCHED_LEARNED) return;
  unsigned const max_edges = p.persist_edges < kPersistEdges ? p.persist_edges : kPersistEdges;
  if (max_edges == 0) return;
  unsigned iter = p.ctl->iter;
  if (iter < 2) return;
  unsigned long long const nw = p.nwords;
  unsigned long long const n = p.n;
  unsigned const max_iter = p.ctl->max_iter;
  // The state the first iteration consumes; later iterations carry it in
  // registers: the variant, the queue iteration - 1 appended (length c1,
  // out-edges e1), the length c2 of the one iteration - 2 appended, and the
  // near-far state (threshold, the previous one, pile size, least pile
  // value, current pile).
  unsigned variant = p.ctl->next_variant;
  unsigned c1 = p.ctl->q_count;
  unsigned long long e1 = p.ctl->q_edges;
  unsigned c2 = p.stats[iter - 2].count;
  unsigned long long T = ~0ull, Told = 0, fvmin = ~0ull;
  unsigned fn = 0, fsel = 0;
  if constexpr (FAR) {
    T = p.ctl->thresh;
    Told = p.ctl->thresh_prev;
    fvmin = p.ctl->far_vmin;
    fn = p.ctl->far_n;
    fsel = p.ctl->far_sel & 1u;
  }
  if (!persist_fits<FAR, vT>(iter, variant, c1, e1, fn, max_edges)) return;
  for (;;) {
    bool const split = FAR && variant == V_SPLIT;
    unsigned const in_count = split ? fn : c1;
    const unsigned* in_q = p.q + (iter & 1) * n;
    unsigned* out_q = p.q + ((iter + 1) & 1) * n;
    const unsigned* in_qe = p.qe + (iter & 1) * (n + 1);
    unsigned* out_qe = p.qe + ((iter + 1) & 1) * (n + 1);
    unsigned long long* out_bm = p.bm + ((iter + 1) % 3) * nw;
    unsigned long long* clr_bm = p.bm + ((iter + 2) % 3) * nw;
    // 1. clear the bits of the queue iteration - 2 appended (it is the
    // queue slot this iteration appends to; after a large step_kernel
    // iteration it can hold up to n entries, and then the whole slot is
    // cleared instead: coalesced stores of nw words beat c2 scattered ones)
    if (c2 > nw / 4) {
      for (unsigned long long w = t; w < nw; w += kPersistBlock) clr_bm[w] = 0ull;
    } else {
      unsigned i = t;
      for (; i + 3 * kPersistBlock < c2; i += 4 * kPersistBlock) {
        unsigned a[4];
#pragma unroll
        for (unsigned k = 0; k < 4; ++k) a[k] = out_q[i + k * kPersistBlock];
#pragma unroll
        for (unsigned k = 0; k < 4; ++k) clr_bm[a[k] >> 6] = 0ull;
      }
      for (; i < c2; i += kPersistBlock) clr_bm[out_q[i] >> 6] = 0ull;
    }
    // 2. stage the queue (every load first; the sentinel offset is e1)
    if (!split) {
      constexpr unsigned kPer = kItems / kPersistBlock;
      unsigned u[kPer], o[kPer], b[kPer];
      vT v[kPer];
#pragma unroll
      for (unsigned k = 0; k < kPer; ++k) {
        unsigned const i = t + k * kPersistBlock;
        u[k] = i < in_count ? in_q[i] : 0u;
        o[k] = i < in_count ? in_qe[i] : 0u;
      }
#pragma unroll
      for (unsigned k = 0; k < kPer; ++k) {
        unsigned const i = t + k * kPersistBlock;
        v[k] = vT{};
        b[k] = 0u;
        if (i < in_count) {
          v[k] = load_agent(p.V + u[k]);
          b[k] = p.oR[u[k]];
        }
      }
#pragma unroll
      for (unsigned k = 0; k < kPer; ++k) {
        unsigned const i = t + k * kPersistBlock;
        if (i < in_count) {
          sm.off[i] = o[k];
          sm.val[i] = v[k];
          sm.base[i] = b[k];
        }
      }
    }
    if (t == 0) {
      if (!split) sm.off[in_count] = static_cast<unsigned>(e1);
      sm.qe = 0ull;
      sm.maxdeg = 0u;
      sm.far_add = 0u;
      sm.vmin = ~0ull;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    unsigned wmax = 0;
    unsigned long long lvmin = ~0ull;   // near-far: least far value this lane saw
    unsigned long long work = 0;
    if (!split) {
      // 3. the edges, kEPT rounds of 1,024 at a time: every load of a pass
      // is issued before any is used
      unsigned const E = static_cast<unsigned>(e1);
      work = E;
      unsigned* const far_q = FAR ? p.fq + static_cast<unsigned long long>(fsel) * n : nullptr;
      unsigned const rounds = (E + kPersistBlock - 1) / kPersistBlock;   // uniform
      for (unsigned r = 0; r < rounds; r += kEPT) {
        unsigned own[kEPT], idx[kEPT], dst[kEPT], deg[kEPT];
        vT cand[kEPT];
        bool ok[kEPT];
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          unsigned const e = t + (r + k) * kPersistBlock;
          ok[k] = e < E;
          own[k] = 0u;
          idx[k] = 0u;
          if (ok[k]) {
            own[k] = persist_owner<kItems>(sm.off, in_count, e);
            idx[k] = sm.base[own[k]] + (e - sm.off[own[k]]);
          }
        }
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) dst[k] = ok[k] ? p.oC[idx[k]] : 0u;
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          cand[k] = vT{};
          if (ok[k]) {
            p.fn(vs(sm.slot[t]), sm.val[own[k]], p.oE + idx[k]);
            cand[k] = sm.slot[t];
          }
        }
        // One returning atomic min per edge (a filtering read first would
        // cost this one compute unit a second scattered access per edge);
        // the degree of a destination it lowered, then its bit.
        bool imp[kEPT];
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          imp[k] = false;
          if (ok[k]) {
            vT const old = atomic_min_any(p.V + dst[k], cand[k]);
            imp[k] = cand[k] < old;
          }
        }
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k)
          deg[k] = imp[k] ? p.oR[dst[k] + 1] - p.oR[dst[k]] : 0u;
        bool newly[kEPT], far[kEPT];
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          newly[k] = false;
          far[k] = false;
          if (deg[k]) {
            unsigned long long const bit = 1ull << (dst[k] & 63);
            bool near = true;
            if constexpr (FAR) {
              unsigned long long const x = static_cast<unsigned long long>(cand[k]);
              near = x < T;
              if (!near) {
                lvmin = x < lvmin ? x : lvmin;
                far[k] = !(atomicOr(p.fbm + (dst[k] >> 6), bit) & bit);
              }
            }
            if (near) newly[k] = !(atomicOr(out_bm + (dst[k] >> 6), bit) & bit);
          }
        }
        // appends: one LDS atomic per wave and round (per list)
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          persist_append(sm, newly[k], dst[k], deg[k], out_q, out_qe, wmax);
          if constexpr (FAR) persist_far_append(sm, far[k], dst[k], far_q, fn);
        }
      }
    } else if constexpr (FAR) {
      // 3'. split: every pile entry is dropped, moved to the queue or kept
      // on the other pile (far bits of entries leaving the pile cleared)
      const unsigned* const src = p.fq + static_cast<unsigned long long>(fsel) * n;
      unsigned* const keep_q = p.fq + static_cast<unsigned long long>(fsel ^ 1u) * n;
      unsigned const rounds = (in_count + kPersistBlock - 1) / kPersistBlock;   // uniform
      work = in_count;
      for (unsigned r = 0; r < rounds; r += kEPT) {
        unsigned u[kEPT], deg[kEPT];
        unsigned long long x[kEPT];
        bool ok[kEPT];
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          unsigned const i = t + (r + k) * kPersistBlock;
          ok[k] = i < in_count;
          u[k] = ok[k] ? src[i] : 0u;
        }
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k)
          x[k] = ok[k] ? static_cast<unsigned long long>(load_agent(p.V + u[k])) : ~0ull;
        bool keep[kEPT], move[kEPT];
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          keep[k] = ok[k] && x[k] >= T;
          move[k] = ok[k] && !keep[k] && x[k] >= Told;
          if (keep[k]) lvmin = x[k] < lvmin ? x[k] : lvmin;
          if (ok[k] && !keep[k]) atomicAnd(p.fbm + (u[k] >> 6), ~(1ull << (u[k] & 63)));
          // pile entries have out-edges (a far bit is claimed only then)
          deg[k] = move[k] ? p.oR[u[k] + 1] - p.oR[u[k]] : 0u;
        }
        bool newly[kEPT];
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          newly[k] = false;
          if (move[k]) {
            unsigned long long const bit = 1ull << (u[k] & 63);
            newly[k] = !(atomicOr(out_bm + (u[k] >> 6), bit) & bit);
          }
        }
#pragma unroll
        for (unsigned k = 0; k < kEPT; ++k) {
          persist_append(sm, newly[k], u[k], deg[k], out_q, out_qe, wmax);
          persist_far_append(sm, keep[k], u[k], keep_q, 0u);
        }
      }
    }
    wmax = wave_max(wmax);
    if (lane == 0 && wmax) atomicMax(&sm.maxdeg, wmax);
    if constexpr (FAR) {
      lvmin = wave_min_u64(lvmin);
      if (lane == 0 && lvmin != ~0ull) atomicMin(&sm.vmin, lvmin);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // 4. close the next queue, the statistics, the next variant
    unsigned long long const qe = sm.qe;
    unsigned const count = qe_count(qe);
    unsigned const sdeg = qe_edges(qe);
    unsigned const maxdeg = sm.maxdeg;
    // near-far: the pile after this iteration
    unsigned fn_next = fn;
    unsigned long long fvmin_next = fvmin;
    bool split_next = false;
    if constexpr (FAR) {
      if (split) {
        fn_next = sm.far_add;
        fvmin_next = sm.vmin;   // exact: the least value of the entries kept
      } else {
        fn_next = fn + sm.far_add;
        fvmin_next = sm.vmin < fvmin ? sm.vmin : fvmin;
      }
      split_next = count == 0 && fn_next != 0;
    }
    bool const stop = (count == 0 && !split_next) || iter + 1 >= max_iter;
    unsigned long long T_next = T;
    if (split_next) {
      unsigned long long const tt = fvmin_next + p.sched.delta;
      T_next = tt < fvmin_next ? ~0ull : tt;
    }
    if (t == 0) {
      out_qe[count] = sdeg;
      step_stats* st = p.stats + iter;
      st->count = count;
      st->variant = split ? static_cast<unsigned>(V_SPLIT) : static_cast<unsigned>(V_PERSIST);
      st->max_outdeg = maxdeg;
      st->in_count = in_count;
      st->sum_outdeg = sdeg;
      st->work_edges = work;
      st->gathers = 0;
      st->active = in_count;
      st->qe = qe;
    }
    if (wib == 0) {
      unsigned next = V_CLASSIC;
      if (!stop) {
        learned_in li;
        li.scratch = sm.mlp;
        if (p.sched.mode == SCHED_LEARNED) {
          li.cache_key = p.ctl->cache_key;
          li.cache_variant = p.ctl->cache_variant;
        }
        next = split_next ? static_cast<unsigned>(V_SPLIT)
                          : choose_variant<true>(p, count, sdeg, maxdeg, iter + 1, li);
      }
      if (lane == 0) {
        p.ctl->next_variant = next;
        p.ctl->iter = iter + 1;
        p.ctl->small_next = stop         ? 0u
                            : split_next ? (fn_next <= kSmallWork ? 1u : 0u)
                                         : grid_level(next, count, sdeg);
        p.ctl->q_count = count;
        p.ctl->q_edges = sdeg;
        if constexpr (FAR) {
          if (split) {
            p.ctl->far_sel = fsel ^ 1u;
            p.ctl->far_next = 0;
          }
          p.ctl->far_n = fn_next;
          p.ctl->far_vmin = fvmin_next;
          if (split_next) {
            p.ctl->thresh_prev = T;
            p.ctl->thresh = T_next;
          }
        }
        if (stop) {
          p.ctl->done = 1;
          if (count != 0 || split_next) p.ctl->overflow = 1;
        }
        sm.next = next;
        sm.go = (!stop && persist_fits<FAR, vT>(iter + 1, next, count, sdeg, fn_next, max_edges))
                    ? 1u
                    : 0u;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!sm.go) return;
    variant = sm.next;
    c2 = c1;
    c1 = count;
    e1 = sdeg;
    if constexpr (FAR) {
      if (split) fsel ^= 1u;
      fn = fn_next;
      fvmin = fvmin_next;
      if (split_next) {
        Told = T;
        T = T_next;
      }
    }
    ++iter;
    // sm.next and sm.go are rewritten only after the next iteration's first
    // barrier, which every thread reaches after reading them here
  }
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_PERSIST_KERNEL_HPP_
