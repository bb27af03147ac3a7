// bitbfs_kernel.hpp -- bit-parallel BFS from up to 256 sources at once.
//
// Every vertex carries 256-bit masks (4 x u64, 32 B): VIS (sources that
// have reached it) and, per iteration, F (sources that reached it in the
// previous iteration).  One launch is one BFS level d, pull over the in-edge
// CSR, level-synchronous:
//   NF[v] = (OR over in-edges (u, v) of F_in[u]) & ~VIS[v]
//   VIS[v] |= NF[v];  F_out[v] = NF[v];  level[v][k] = d + 1 for k in NF[v].
// A one-bit-per-vertex summary A_in (F_in[u] != 0) lets a lane skip the
// 32-B mask load of an inactive source; it fits in L2 at RMAT-22 (512 KB).
// A second one, DONE (VIS[v] holds every source), lets the lanes of a
// destination that every source has reached skip theirs: R-MAT-22 with 256
// sources, 11.3 -> 10.6 ms (the two dense levels gather 62M and 46M masks;
// the level after them mostly lands on done vertices).
//
// Work items are the fixed edge ranges of the K-wide kernel
// (batch_kernel.hpp).  Per 64-edge chunk each lane holds one edge's source
// mask; a segmented OR over the lanes of each destination (runs are
// contiguous: the in-edge list is sorted by destination) leaves the
// destination's OR in the run's last lane; a run that continues into the
// next chunk is carried in registers.  A destination whose edges all lie in
// the item is owned by the wave (plain loads and stores); one cut by an
// item boundary uses atomicOr, and two waves may then both see a bit as new
// and both write the same level, which is harmless.
//
// Levels are kept as bytes while the run goes (L8, n x 256 B, written only
// for the pairs reached): the whole wave writes each new row, lane l
// covering sources 4l .. 4l+3 with byte stores, so a level costs 1 B of HBM
// writes instead of a masked 4-B store into a 1-KB row.  Levels 255 and
// beyond (runs deeper than 254 levels) are stored straight into the u32
// output L, their byte set to 0xFF.  After the run, bitbfs_expand_kernel
// writes every u32 row of L once, whole and coalesced, from its byte row
// and VIS (no VIS bit = unreached, whatever the byte holds).
//
// The frontier slot an iteration will write is cleared two iterations
// ahead, but only where it was written: the one-bit summary of that slot
// (A_clr) names the vertices whose masks are non-zero, and one wave per
// 64-vertex word clears their 32-B masks.
//
// A level whose frontier has few out-edges runs top-down instead (push):
// a pull walks every in-edge whatever the frontier (about 0.75 ms a level
// at R-MAT-22, five of its nine levels reaching almost nothing), while a
// push walks only the frontier's out-edges.  One wave per 64-vertex word
// of the frontier summary walks each of its frontier vertices' out-edges
// (the whole workgroup walks a long list together), 64 at a time: for
// frontier vertex u and edge (u, v), new = F_in[u] & ~VIS[v], then a returning
// atomicOr per non-zero VIS word keeps exactly the bits this edge adds (each
// (vertex, source) pair is claimed once), which go into F_out[v] by
// atomicOr and into the byte levels.  A returning atomicOr on the summary
// bit counts each updated vertex once, with its out-degree, so the last
// workgroup can pick the next level's direction: push when the new
// frontier's out-edges * push_div < m and none has more than
// kBitPushMaxDeg (a workgroup walks one vertex's list alone).
#ifndef KITES_DETAIL_BITBFS_KERNEL_HPP_
#define KITES_DETAIL_BITBFS_KERNEL_HPP_

#include <hip/hip_runtime.h>

#include "batch_kernel.hpp"
#include "wave.hpp"

namespace kites {
namespace detail {

constexpr unsigned kBitSources = 256;
constexpr unsigned kBitWords = kBitSources / 64;

struct bitbfs_stats {
  unsigned updated;                // vertices that gained sources
  unsigned push;                   // 1: the level ran top-down
  unsigned long long active_edges; // edges from an active source into a vertex not yet done
  unsigned long long new_pairs;    // (vertex, source) pairs reached
  unsigned long long out_edges;    // out-edges of the vertices updated (the next frontier)
  unsigned max_out;                // largest out-degree among them
  unsigned pad;
};

// Largest out-degree a top-down level's frontier may hold.
constexpr unsigned kBitPushMaxDeg = 4096;
// Default direction rule: top-down when the frontier's out-edges are
// below m / kBitPushDiv.
constexpr unsigned kBitPushDiv = 4;
// A top-down frontier vertex with more out-edges than this is walked by
// its workgroup's waves together (at most kBitHeavyCap per workgroup).
constexpr unsigned kBitPushHeavy = 256;
constexpr unsigned kBitHeavyCap = 256;

struct bitbfs_params {
  unsigned n;
  unsigned m;
  const unsigned* C;
  const unsigned* nzi;
  const unsigned* nzR;
  unsigned nnz;
  const unsigned* istart;       // non-empty-vertex position of every 64-edge chunk
  unsigned nitems;
  unsigned epi;                 // edges per item, a multiple of 64
  unsigned long long* VIS;      // n x 4
  unsigned long long* F;        // 3 slots of n x 4
  unsigned long long* A;        // 3 slots of aw words
  unsigned aw;
  // One bit per vertex whose VIS holds every source (`full`): no in-edge
  // can add to it, so its in-edges skip the source mask loads.
  unsigned long long* DONE;     // aw words
  unsigned long long full[kBitWords];
  unsigned* L;                  // n x 256 levels (u32 output)
  unsigned char* L8;            // n x 256 levels below 255 (0xFF = none yet)
  batch_ctl* ctl;
  bitbfs_stats* stats;
  const unsigned* oR;           // out-edge CSR (top-down levels)
  const unsigned* oC;
  unsigned push_div;            // top-down when frontier out-edges * push_div < m (0: never)
};

__device__ __forceinline__ unsigned long long shfl_up64(unsigned long long x, unsigned s) {
  unsigned const lo = __shfl_up(static_cast<unsigned>(x), s, 64);
  unsigned const hi = __shfl_up(static_cast<unsigned>(x >> 32), s, 64);
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

static __global__ void bitbfs_init_kernel(unsigned long long* __restrict__ VIS, unsigned long long nvis,
                                   unsigned long long* __restrict__ F, unsigned long long nf,
                                   unsigned long long* __restrict__ A, unsigned long long na) {
  unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long const st = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
  for (unsigned long long j = i; j < nvis; j += st) VIS[j] = 0ull;
  for (unsigned long long j = i; j < nf; j += st) F[j] = 0ull;
  for (unsigned long long j = i; j < na; j += st) A[j] = 0ull;
}

// Whether a level whose frontier has `out_edges` out-edges, none of its
// vertices more than `max_out`, runs top-down.
__device__ __forceinline__ bool bitbfs_push_next(bitbfs_params const& p,
                                                 unsigned long long out_edges, unsigned max_out) {
  return p.push_div != 0u && out_edges * p.push_div < p.m && max_out <= kBitPushMaxDeg;
}

// Byte levels of the rows the lanes with `any` set gained (the bits of nf
// in row `key`), written by the whole wave: lane l covers sources 4l ..
// 4l + 3 of each row in turn.  All 64 lanes, uniform control flow.
__device__ __forceinline__ void bitbfs_write_levels(bitbfs_params const& p, unsigned level,
                                                    bool any, unsigned key,
                                                    unsigned long long const* nf) {
  unsigned const lane = lane_id();
  unsigned long long newm = ballot64(any);
  while (newm) {
    unsigned const j = __builtin_ctzll(newm);
    newm &= newm - 1;
    unsigned const d = readlane(key, j);
    unsigned long long const w0 = readlane64(nf[0], j), w1 = readlane64(nf[1], j);
    unsigned long long const w2 = readlane64(nf[2], j), w3 = readlane64(nf[3], j);
    unsigned const wi = lane >> 4;   // sources 4 lane .. 4 lane + 3 sit in word lane / 16
    unsigned long long const word = wi == 0 ? w0 : (wi == 1 ? w1 : (wi == 2 ? w2 : w3));
    unsigned const bits = static_cast<unsigned>(word >> ((4u * lane) & 63u)) & 0xfu;
    if (bits) {
      unsigned long long const at = static_cast<unsigned long long>(d) * kBitSources + 4u * lane;
      if (level < 255u) {
        unsigned char* row = p.L8 + at;
#pragma unroll
        for (unsigned q = 0; q < 4; ++q)
          if ((bits >> q) & 1u) row[q] = static_cast<unsigned char>(level);
      } else {
        unsigned* row = p.L + at;
        unsigned char* brow = p.L8 + at;
#pragma unroll
        for (unsigned q = 0; q < 4; ++q)
          if ((bits >> q) & 1u) {
            row[q] = level;
            brow[q] = 0xFFu;   // the level is in L
          }
      }
    }
  }
}

static __global__ void bitbfs_seed_kernel(bitbfs_params p, const unsigned* __restrict__ src, unsigned K,
                                   unsigned max_iter) {
  __shared__ unsigned long long s_out;
  __shared__ unsigned s_max;
  unsigned const k = threadIdx.x;
  if (k == 0) {
    s_out = 0ull;
    s_max = 0u;
  }
  __syncthreads();
  if (k < K) {
    unsigned const s0 = src[k];
    unsigned const d = p.oR[s0 + 1] - p.oR[s0];
    atomicAdd(&s_out, static_cast<unsigned long long>(d));
    atomicMax(&s_max, d);
  }
  __syncthreads();
  if (k < K) {
    unsigned const s = src[k];
    unsigned long long const bit = 1ull << (k & 63);
    atomicOr(p.VIS + static_cast<unsigned long long>(s) * kBitWords + (k >> 6), bit);
    atomicOr(p.F + static_cast<unsigned long long>(s) * kBitWords + (k >> 6), bit);
    atomicOr(p.A + (s >> 6), 1ull << (s & 63));
    p.L8[static_cast<unsigned long long>(s) * kBitSources + k] = 0u;
  }
  if (k == 0) {
    p.ctl->iter = 0;
    p.ctl->done = K == 0;
    p.ctl->max_iter = max_iter;
    p.ctl->overflow = 0;
    p.ctl->push = bitbfs_push_next(p, s_out, s_max) ? 1u : 0u;
  }
  clear_ticket_board(p.ctl->board.w, k, blockDim.x);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void bitbfs_step_kernel(bitbfs_params p) {
  constexpr unsigned kW = BLOCK / 64;
  __shared__ unsigned heads[kW][64];
  __shared__ unsigned is_last;
  __shared__ unsigned s_heavy[kBitHeavyCap];   // top-down: vertices the workgroup walks together
  __shared__ unsigned s_nheavy;
  __shared__ unsigned s_wupd[kW], s_wmax[kW];            // per-wave statistics, summed by thread 0
  __shared__ unsigned long long s_wact[kW], s_wpairs[kW], s_wout[kW];
  if (p.ctl->done) return;
  unsigned const iter = p.ctl->iter;
  unsigned const n = p.n;
  unsigned long long const slot = static_cast<unsigned long long>(n) * kBitWords;
  unsigned long long* F_in = p.F + (iter % 3) * slot;
  unsigned long long* F_out = p.F + ((iter + 1) % 3) * slot;
  unsigned long long* F_clr = p.F + ((iter + 2) % 3) * slot;
  unsigned long long const* A_in = p.A + static_cast<unsigned long long>(iter % 3) * p.aw;
  unsigned long long* A_out = p.A + static_cast<unsigned long long>((iter + 1) % 3) * p.aw;
  unsigned long long* A_clr = p.A + static_cast<unsigned long long>((iter + 2) % 3) * p.aw;
  unsigned const t = threadIdx.x, lane = lane_id(), wib = t >> 6;
  {
    // Clear the slot iteration iter + 1 writes, where it was written (the
    // vertices its summary marks), then the summary itself; one wave per
    // summary word, lane l taking vertex 64 w + l.
    unsigned const gw = blockIdx.x * kW + wib, nwv = gridDim.x * kW;
    for (unsigned w = gw; w < p.aw; w += nwv) {
      unsigned long long const word = A_clr[w];
      if (!word) continue;
      if ((word >> lane) & 1ull) {
        ulonglong2* fp = reinterpret_cast<ulonglong2*>(
            F_clr + (static_cast<unsigned long long>(w) * 64u + lane) * kBitWords);
        fp[0] = make_ulonglong2(0ull, 0ull);
        fp[1] = make_ulonglong2(0ull, 0ull);
      }
      if (lane == 0) A_clr[w] = 0ull;
    }
  }
  unsigned const level = iter + 1;
  unsigned upd = 0, max_next = 0;
  unsigned long long act_edges = 0, pairs = 0, out_next = 0;
  unsigned const nw = gridDim.x * kW;
  bool const push = p.ctl->push != 0u;
  if (push) {
    // Top-down.  Walks frontier vertex u's out-edges 64 at a time from
    // edge oR[u] + first, `step` edges apart (one wave: 64; the whole
    // workgroup: 64 * kW); uniform per wave.
    auto walk = [&](unsigned u, unsigned first, unsigned step) {
      unsigned long long const* fu = F_in + static_cast<unsigned long long>(u) * kBitWords;
      unsigned long long mk[kBitWords];
#pragma unroll
      for (unsigned q = 0; q < kBitWords; ++q) mk[q] = fu[q];
      unsigned const eb = p.oR[u], ee = p.oR[u + 1];
      for (unsigned b = eb + first; b < ee; b += step) {
        unsigned const e = b + lane;
        bool const valid = e < ee;
        unsigned const v = valid ? p.oC[e] : 0u;
        bool const act = valid && !((p.DONE[v >> 6] >> (v & 63)) & 1ull);
        act_edges += act ? 1u : 0u;
        unsigned long long nf[kBitWords] = {0ull, 0ull, 0ull, 0ull};
        bool any = false;
        if (act) {
          unsigned long long* vis = p.VIS + static_cast<unsigned long long>(v) * kBitWords;
          ulonglong2 const* vp2 = reinterpret_cast<ulonglong2 const*>(vis);
          ulonglong2 const a = vp2[0], c = vp2[1];
          unsigned long long const vo[kBitWords] = {a.x, a.y, c.x, c.y};
          unsigned long long vn[kBitWords];
          unsigned long long* fo = F_out + static_cast<unsigned long long>(v) * kBitWords;
#pragma unroll
          for (unsigned q = 0; q < kBitWords; ++q) {
            unsigned long long const want = mk[q] & ~vo[q];
            vn[q] = vo[q];
            if (want) {
              // returning: keeps only the bits no other edge claimed first
              unsigned long long const old = atomicOr(vis + q, want);
              nf[q] = want & ~old;
              vn[q] = old | want;
              if (nf[q]) atomicOr(fo + q, nf[q]);
            }
          }
          any = (nf[0] | nf[1] | nf[2] | nf[3]) != 0ull;
          if (any) {
            if (vn[0] == p.full[0] && vn[1] == p.full[1] && vn[2] == p.full[2] &&
                vn[3] == p.full[3])
              atomicOr(p.DONE + (v >> 6), 1ull << (v & 63));
            unsigned long long const bit = 1ull << (v & 63);
            // the edge that sets v's summary bit counts v (once)
            if (!(atomicOr(A_out + (v >> 6), bit) & bit)) {
              ++upd;
              unsigned const od = p.oR[v + 1] - p.oR[v];
              out_next += od;
              max_next = od > max_next ? od : max_next;
            }
            pairs += __popcll(nf[0]) + __popcll(nf[1]) + __popcll(nf[2]) + __popcll(nf[3]);
          }
        }
        bitbfs_write_levels(p, level, any, v, nf);
      }
    };
    // One wave per word of the frontier summary walks its light vertices
    // alone; a vertex with more than kBitPushHeavy out-edges is listed for
    // the whole workgroup, which walks the list's vertices together after
    // every wave is done with its words (a full list: the wave walks it).
    if (t == 0) s_nheavy = 0u;
    __syncthreads();
    for (unsigned w = blockIdx.x * kW + wib; w < p.aw; w += nw) {
      unsigned long long word = A_in[w];
      while (word) {
        unsigned const u = w * 64u + static_cast<unsigned>(__builtin_ctzll(word));
        word &= word - 1ull;
        if (p.oR[u + 1] - p.oR[u] > kBitPushHeavy) {
          unsigned slot = 0;
          if (lane == 0) slot = atomicAdd(&s_nheavy, 1u);
          slot = readlane(slot, 0);
          if (slot < kBitHeavyCap) {
            if (lane == 0) s_heavy[slot] = u;
            continue;
          }
        }
        walk(u, 0u, 64u);
      }
    }
    __syncthreads();
    unsigned const nh = s_nheavy < kBitHeavyCap ? s_nheavy : kBitHeavyCap;
    for (unsigned i = 0; i < nh; ++i) walk(s_heavy[i], 64u * wib, 64u * kW);
  }
  for (unsigned k = push ? p.nitems : blockIdx.x * kW + wib; k < p.nitems; k += nw) {
    unsigned const e0 = k * p.epi;
    unsigned const e1 = min(e0 + p.epi, p.m);
    unsigned vp = p.istart[(k * p.epi) >> 6];   // chunk starts, one per 64 edges
    unsigned cur = 0xffffffffu;          // carried destination (uniform)
    unsigned long long cacc[kBitWords] = {0ull, 0ull, 0ull, 0ull};
    // VIS of the carried destination, read when its run began to carry, and
    // whether the run's OR so far covers every source it lacks (then its
    // remaining in-edges can add nothing: their masks are not loaded).
    unsigned long long cvis[kBitWords] = {0ull, 0ull, 0ull, 0ull};
    bool csat = false;
    for (unsigned b = e0; b < e1; b += 64) {
      unsigned const e = b + lane;
      bool const valid = e < e1;
      unsigned const u = valid ? p.C[e] : 0u;
      bool act = valid && ((A_in[u >> 6] >> (u & 63)) & 1ull);
      unsigned const pos = vp + lane;
      unsigned const end = pos < p.nnz ? p.nzR[pos + 1] : 0xffffffffu;
      unsigned const start = pos < p.nnz ? p.nzR[pos] : 0xffffffffu;
      unsigned const vid = pos < p.nnz ? p.nzi[pos] : 0u;
      heads[wib][lane] = 0;
      wave_lds_sync();
      if (end > b && end - b < 64u) atomicMax(&heads[wib][end - b], lane + 1);
      wave_lds_sync();
      unsigned const off = wave_incl_max(heads[wib][lane]);
      unsigned const key = valid ? __shfl(vid, static_cast<int>(off), 64) : 0xffffffffu;
      unsigned const vend = __shfl(end, static_cast<int>(off), 64);
      bool const vcut = __shfl(static_cast<unsigned>(start < e0 || end > e1), static_cast<int>(off), 64) != 0u;
      unsigned const adv = __popcll(ballot64(end <= b + 64u));
      // an edge into a vertex every source has reached adds nothing: its
      // source mask is not loaded
      if (act && ((p.DONE[key >> 6] >> (key & 63)) & 1ull)) act = false;
      if (act && csat && key == cur) act = false;
      unsigned long long x[kBitWords] = {0ull, 0ull, 0ull, 0ull};
      if (act) {
        ulonglong2 const* fp = reinterpret_cast<ulonglong2 const*>(F_in + static_cast<unsigned long long>(u) * kBitWords);
        ulonglong2 const a = fp[0], c = fp[1];
        x[0] = a.x; x[1] = a.y; x[2] = c.x; x[3] = c.y;
      }
      act_edges += act ? 1u : 0u;
      // carried run continues in this chunk's first run
      if (key == cur) {
#pragma unroll
        for (unsigned w = 0; w < kBitWords; ++w) x[w] |= cacc[w];
      }
      // segmented inclusive OR towards the last lane of each run
#pragma unroll
      for (unsigned s = 1; s < 64; s <<= 1) {
        unsigned const k2 = __shfl_up(key, s, 64);
        bool const same = lane >= s && k2 == key;
#pragma unroll
        for (unsigned w = 0; w < kBitWords; ++w) {
          unsigned long long const y = shfl_up64(x[w], s);
          if (same) x[w] |= y;
        }
      }
      unsigned const knext = __shfl_down(key, 1, 64);
      bool const last = valid && (lane == 63 || knext != key);
      // the chunk's final run continues if its vertex ends past this chunk
      bool const carry = last && vend > b + 64u && e1 > b + 64u;
      unsigned long long const cm = ballot64(carry);
      // flush every other completed run
      bool const fl = last && !carry;
      unsigned long long nf[kBitWords] = {0ull, 0ull, 0ull, 0ull};
      bool any = false;
      if (fl) {
        unsigned long long* vis = p.VIS + static_cast<unsigned long long>(key) * kBitWords;
        unsigned long long vn[kBitWords];   // VIS after this flush (as far as seen)
        if (!vcut) {
#pragma unroll
          for (unsigned w = 0; w < kBitWords; ++w) {
            unsigned long long const vo = vis[w];
            nf[w] = x[w] & ~vo;
            vn[w] = vo | nf[w];
          }
          any = (nf[0] | nf[1] | nf[2] | nf[3]) != 0ull;
          if (any) {
            unsigned long long* fo = F_out + static_cast<unsigned long long>(key) * kBitWords;
#pragma unroll
            for (unsigned w = 0; w < kBitWords; ++w) {
              vis[w] = vn[w];
              fo[w] = nf[w];
            }
          }
        } else {
#pragma unroll
          for (unsigned w = 0; w < kBitWords; ++w) {
            unsigned long long const vo = vis[w];
            nf[w] = x[w] & ~vo;
            vn[w] = vo | nf[w];
          }
          any = (nf[0] | nf[1] | nf[2] | nf[3]) != 0ull;
          if (any) {
            unsigned long long* fo = F_out + static_cast<unsigned long long>(key) * kBitWords;
#pragma unroll
            for (unsigned w = 0; w < kBitWords; ++w) {
              if (nf[w]) {
                vn[w] = atomicOr(vis + w, nf[w]) | nf[w];
                atomicOr(fo + w, nf[w]);
              }
            }
          }
        }
        if (any) {
          // a row that now holds every source is done (a cut row whose
          // writers each miss the other's bits stays unmarked: it then
          // only keeps its loads)
          if (vn[0] == p.full[0] && vn[1] == p.full[1] && vn[2] == p.full[2] &&
              vn[3] == p.full[3])
            atomicOr(p.DONE + (key >> 6), 1ull << (key & 63));
          atomicOr(A_out + (key >> 6), 1ull << (key & 63));
          ++upd;
          pairs += __popcll(nf[0]) + __popcll(nf[1]) + __popcll(nf[2]) + __popcll(nf[3]);
          // (a row cut by an item boundary may count twice: a direction
          // hint only)
          unsigned const od = p.oR[key + 1] - p.oR[key];
          out_next += od;
          max_next = od > max_next ? od : max_next;
        }
      }
      // levels of the new rows, written by the whole wave
      bitbfs_write_levels(p, level, any, key, nf);
      // carry the final run into the next chunk
      if (cm) {
        unsigned const j = 63u - __builtin_clzll(cm);
        unsigned const nk = readlane(key, j);
#pragma unroll
        for (unsigned w = 0; w < kBitWords; ++w) cacc[w] = readlane64(x[w], j);
        if (nk != cur) {
          cur = nk;
          ulonglong2 const* vp2 =
              reinterpret_cast<ulonglong2 const*>(p.VIS + static_cast<unsigned long long>(cur) * kBitWords);
          ulonglong2 const a = vp2[0], c = vp2[1];
          // (one address for every lane: held as wave-uniform values)
          cvis[0] = readlane64(a.x, 0);
          cvis[1] = readlane64(a.y, 0);
          cvis[2] = readlane64(c.x, 0);
          cvis[3] = readlane64(c.y, 0);
        }
        csat = (cacc[0] | cvis[0]) == p.full[0] && (cacc[1] | cvis[1]) == p.full[1] &&
               (cacc[2] | cvis[2]) == p.full[2] && (cacc[3] | cvis[3]) == p.full[3];
      } else {
        cur = 0xffffffffu;
        csat = false;
#pragma unroll
        for (unsigned w = 0; w < kBitWords; ++w) cacc[w] = 0ull;
      }
      vp += adv;
    }
  }

  // Each updated row was counted by the lane that flushed it.  The wave
  // sums go to LDS and thread 0 adds the workgroup's totals: one set of
  // atomics per workgroup instead of one per wave (the counters share a
  // line, and atomics on one line serialize, about 12 ns each).
  {
    unsigned const ut = wave_sum(upd);
    unsigned long long const a = wave_sum64(act_edges);
    unsigned long long const q = wave_sum64(pairs);
    unsigned long long const o = wave_sum64(out_next);
    unsigned const mo = wave_max(max_next);
    if (lane == 0) {
      s_wupd[wib] = ut;
      s_wmax[wib] = mo;
      s_wact[wib] = a;
      s_wpairs[wib] = q;
      s_wout[wib] = o;
    }
  }
  __syncthreads();
  if (t == 0) {
    unsigned ut = 0, mo = 0;
    unsigned long long a = 0, q = 0, o = 0;
#pragma unroll
    for (unsigned w = 0; w < kW; ++w) {
      ut += s_wupd[w];
      mo = s_wmax[w] > mo ? s_wmax[w] : mo;
      a += s_wact[w];
      q += s_wpairs[w];
      o += s_wout[w];
    }
    if (ut) keep_live(atomicAdd(&p.stats[iter].updated, ut));
    if (a) keep_live(atomicAdd(&p.stats[iter].active_edges, a));
    if (q) keep_live(atomicAdd(&p.stats[iter].new_pairs, q));
    if (o) keep_live(atomicAdd(&p.stats[iter].out_edges, o));
    if (mo) keep_live(atomicMax(&p.stats[iter].max_out, mo));
    __builtin_amdgcn_s_waitcnt(0);
    is_last = take_ticket(p.ctl->board.w);
  }
  __syncthreads();
  if (!is_last || t != 0) return;
  unsigned const u = __hip_atomic_load(&p.stats[iter].updated, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long const oe = __hip_atomic_load(&p.stats[iter].out_edges, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
  unsigned const mo = __hip_atomic_load(&p.stats[iter].max_out, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
  p.stats[iter].push = push ? 1u : 0u;
  p.ctl->push = bitbfs_push_next(p, oe, mo) ? 1u : 0u;
  p.ctl->iter = iter + 1;
  if (u == 0) p.ctl->done = 1;
  else if (iter + 1 >= p.ctl->max_iter) {
    p.ctl->done = 1;
    p.ctl->overflow = 1;
  }
}

// Writes every u32 level row of L from its byte row: one wave per vertex,
// lane l turning bytes 4l .. 4l+3 into one 16-B store.  A pair whose VIS bit
// is clear was not reached (0xFFFFFFFF; its byte is whatever an earlier run
// left, so the byte rows need no clearing); a reached pair's byte is its
// level, or 0xFF when the level (255 or deeper; only if the run went that
// deep, `deep`) was stored straight into L, which is then kept.
static __global__ void bitbfs_expand_kernel(unsigned* __restrict__ L,
                                            const unsigned* __restrict__ L8w,
                                            const unsigned long long* __restrict__ VIS,
                                            unsigned n, int deep) {
  unsigned const lane = lane_id();
  unsigned const wpb = blockDim.x / 64;
  for (unsigned v = blockIdx.x * wpb + (threadIdx.x >> 6); v < n; v += gridDim.x * wpb) {
    unsigned long long const at = static_cast<unsigned long long>(v) * (kBitSources / 4) + lane;
    unsigned const b = L8w[at];
    unsigned long long const vw = VIS[static_cast<unsigned long long>(v) * kBitWords + (lane >> 4)];
    unsigned const vb = static_cast<unsigned>(vw >> ((4u * lane) & 63u)) & 0xfu;
    unsigned x[4];
    bool keep = false;
#pragma unroll
    for (unsigned q = 0; q < 4; ++q) {
      unsigned const c = (b >> (8u * q)) & 0xffu;
      bool const reached = (vb >> q) & 1u;
      x[q] = reached ? c : 0xFFFFFFFFu;
      keep = keep || (deep && reached && c == 0xffu);
    }
    uint4* out = reinterpret_cast<uint4*>(L + static_cast<unsigned long long>(v) * kBitSources) + lane;
    if (keep) {
      uint4 const old = *out;
      unsigned const o[4] = {old.x, old.y, old.z, old.w};
#pragma unroll
      for (unsigned q = 0; q < 4; ++q)
        if (((vb >> q) & 1u) && x[q] == 0xffu) x[q] = o[q];
    }
    *out = make_uint4(x[0], x[1], x[2], x[3]);
  }
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_BITBFS_KERNEL_HPP_
