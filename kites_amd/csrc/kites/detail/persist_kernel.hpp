// persist_kernel.hpp -- small push frontiers in one workgroup, many
// iterations per launch.
//
// High-diameter graphs (grids, road-like meshes) run thousands of iterations
// whose queues hold a few thousand vertices.  A full-grid launch of
// step_kernel per iteration then pays its dispatch, the completion ticket
// and the next decision -- about 12 us on MI355X -- for ~10^4 edges of work.
// persist_kernel runs such iterations in one workgroup of 1,024 threads that
// synchronizes with barriers only (there is no other workgroup, so nothing
// waits on one), until the queue is empty (the run is done), the schedule
// picks a variant that is not a push, the queue outgrows persist_items()
// entries or kPersistEdges out-edges (the next step_kernel launch takes
// over), or the iteration limit is reached.  The engine launches it first in
// every chunk it replays after a small push frontier; when the coming
// iteration does not fit, it returns at once.
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
// The compute unit's L1 keeps lines across the iterations of one launch and
// atomics update only the L2, so every read of a location that atomics
// write (V, the bitmaps) is an agent-scope atomic load.  Plain stores of
// this workgroup (queues, offsets, bitmap clears) pass through the same L1.
#ifndef KITES_DETAIL_PERSIST_KERNEL_HPP_
#define KITES_DETAIL_PERSIST_KERNEL_HPP_

#include <hip/hip_runtime.h>

#include <type_traits>

#include "step_kernel.hpp"
#include "wave.hpp"

namespace kites {
namespace detail {

constexpr unsigned kPersistBlock = 1024;
constexpr unsigned kEPT = 8;              // edges per thread and pass
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
  unsigned maxdeg;
  unsigned go;                // 1: the next iteration runs here too
  float mlp[MLP_HID + MLP_OUT];
};

// Whether iteration `iter` (variant `v`, queue of `count` entries with `sdeg`
// out-edges) is one persist_kernel runs.
template <class vT>
__device__ __forceinline__ bool persist_fits(unsigned iter, unsigned v, unsigned count,
                                             unsigned long long sdeg, unsigned max_edges) {
  return iter >= 2 && v >= V_PUSH_T && v <= V_PUSH_LB && count <= persist_items<vT>() &&
         sdeg <= max_edges;
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

template <class P>
__global__ __launch_bounds__(kPersistBlock) void persist_kernel(P p) {
  using vT = typename std::remove_reference<decltype(*p.V)>::type;
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
  // The state the first iteration consumes (later iterations carry it in
  // registers).
  unsigned const variant = p.ctl->next_variant;
  unsigned in_count = p.ctl->q_count;
  unsigned prev_count = p.stats[iter - 2].count;
  unsigned long long sdeg_in = p.ctl->q_edges;
  if (!persist_fits<vT>(iter, variant, in_count, sdeg_in, max_edges)) return;
  for (;;) {
    const unsigned* in_q = p.q + (iter & 1) * n;
    unsigned* out_q = p.q + ((iter + 1) & 1) * n;
    const unsigned* in_qe = p.qe + (iter & 1) * (n + 1);
    unsigned* out_qe = p.qe + ((iter + 1) & 1) * (n + 1);
    unsigned long long* out_bm = p.bm + ((iter + 1) % 3) * nw;
    unsigned long long* clr_bm = p.bm + ((iter + 2) % 3) * nw;
    // 1. clear the bits of the queue iteration - 2 appended (it is the
    // queue slot this iteration appends to; after a large step_kernel
    // iteration it can hold up to n entries)
    {
      unsigned i = t;
      for (; i + 3 * kPersistBlock < prev_count; i += 4 * kPersistBlock) {
        unsigned a[4];
#pragma unroll
        for (unsigned k = 0; k < 4; ++k) a[k] = out_q[i + k * kPersistBlock];
#pragma unroll
        for (unsigned k = 0; k < 4; ++k) clr_bm[a[k] >> 6] = 0ull;
      }
      for (; i < prev_count; i += kPersistBlock) clr_bm[out_q[i] >> 6] = 0ull;
    }
    // 2. stage the queue (every load first; the sentinel offset is sdeg_in)
    {
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
      sm.off[in_count] = static_cast<unsigned>(sdeg_in);
      sm.qe = 0ull;
      sm.maxdeg = 0u;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // 3. the edges, kEPT rounds of 1,024 at a time: every load of a pass
    // is issued before any is used
    unsigned const E = static_cast<unsigned>(sdeg_in);
    unsigned const rounds = (E + kPersistBlock - 1) / kPersistBlock;   // uniform
    unsigned wmax = 0;
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
      for (unsigned k = 0; k < kEPT; ++k) deg[k] = imp[k] ? p.oR[dst[k] + 1] - p.oR[dst[k]] : 0u;
      bool newly[kEPT];
#pragma unroll
      for (unsigned k = 0; k < kEPT; ++k) {
        newly[k] = false;
        if (deg[k]) {
          unsigned long long const bit = 1ull << (dst[k] & 63);
          newly[k] = !(atomicOr(out_bm + (dst[k] >> 6), bit) & bit);
        }
      }
      // appends: one LDS atomic per wave and round
#pragma unroll
      for (unsigned k = 0; k < kEPT; ++k) {
        unsigned long long const m = ballot64(newly[k]);
        if (!m) continue;
        unsigned const d = newly[k] ? deg[k] : 0u;
        unsigned const incl = wave_incl_scan(d);
        unsigned const dsum = readlane(incl, 63);
        unsigned long long old = 0;
        if (lane == 0)
          old = atomicAdd(&sm.qe, qe_pack(static_cast<unsigned>(__builtin_popcountll(m)), dsum));
        unsigned const olo = readlane(static_cast<unsigned>(old), 0);
        unsigned const ohi = readlane(static_cast<unsigned>(old >> 32), 0);
        unsigned long long const o64 = (static_cast<unsigned long long>(ohi) << 32) | olo;
        if (newly[k]) {
          unsigned const pos = qe_count(o64) + rank_in(m);
          out_q[pos] = dst[k];
          out_qe[pos] = qe_edges(o64) + incl - d;
        }
        wmax = d > wmax ? d : wmax;
      }
    }
    wmax = wave_max(wmax);
    if (lane == 0 && wmax) atomicMax(&sm.maxdeg, wmax);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // 4. close the next queue, the statistics, the next variant
    unsigned long long const qe = sm.qe;
    unsigned const count = qe_count(qe);
    unsigned const sdeg = qe_edges(qe);
    unsigned const maxdeg = sm.maxdeg;
    bool const stop = count == 0 || iter + 1 >= max_iter;
    if (t == 0) {
      out_qe[count] = sdeg;
      step_stats* st = p.stats + iter;
      st->count = count;
      st->variant = V_PERSIST;
      st->max_outdeg = maxdeg;
      st->in_count = in_count;
      st->sum_outdeg = sdeg;
      st->work_edges = E;
      st->gathers = 0;
      st->active = in_count;
      st->qe = qe;
    }
    if (wib == 0) {
      unsigned next = V_CLASSIC;
      if (!stop) next = choose_variant(p, count, sdeg, maxdeg, iter + 1, sm.mlp);
      if (lane == 0) {
        p.ctl->next_variant = next;
        p.ctl->iter = iter + 1;
        p.ctl->small_next = stop ? 0u : grid_level(next, count, sdeg);
        p.ctl->q_count = count;
        p.ctl->q_edges = sdeg;
        if (stop) {
          p.ctl->done = 1;
          if (count != 0) p.ctl->overflow = 1;
        }
        sm.go = (!stop && persist_fits<vT>(iter + 1, next, count, sdeg, max_edges)) ? 1u : 0u;
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!sm.go) return;
    prev_count = in_count;
    in_count = count;
    sdeg_in = sdeg;
    ++iter;
  }
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_PERSIST_KERNEL_HPP_
