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
// Levels are kept as bytes while the run goes (L8, n x 256 B, 0xFF = not
// reached yet): the whole wave writes each new row, lane l covering sources
// 4l .. 4l+3 with byte stores, so a level costs 1 B of HBM writes instead
// of a masked 4-B store into a 1-KB row.  Levels 255 and beyond (runs
// deeper than 254 levels) are stored straight into the u32 output L.  After
// the run, bitbfs_expand_kernel writes every u32 row of L once, whole and
// coalesced, from its byte row (0xFF and no VIS bit = unreached).
//
// The frontier slot an iteration will write is cleared two iterations
// ahead, but only where it was written: the one-bit summary of that slot
// (A_clr) names the vertices whose masks are non-zero, and one wave per
// 64-vertex word clears their 32-B masks.
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
  unsigned pad;
  unsigned long long active_edges; // in-edges whose source was active
  unsigned long long new_pairs;    // (vertex, source) pairs reached
};

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
};

__device__ __forceinline__ unsigned long long shfl_up64(unsigned long long x, unsigned s) {
  unsigned const lo = __shfl_up(static_cast<unsigned>(x), s, 64);
  unsigned const hi = __shfl_up(static_cast<unsigned>(x >> 32), s, 64);
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

static __global__ void bitbfs_init_kernel(unsigned* __restrict__ L8w, unsigned long long nl8w,
                                   unsigned long long* __restrict__ VIS, unsigned long long nvis,
                                   unsigned long long* __restrict__ F, unsigned long long nf,
                                   unsigned long long* __restrict__ A, unsigned long long na) {
  unsigned long long const i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned long long const st = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
  for (unsigned long long j = i; j < nl8w; j += st) L8w[j] = 0xFFFFFFFFu;
  for (unsigned long long j = i; j < nvis; j += st) VIS[j] = 0ull;
  for (unsigned long long j = i; j < nf; j += st) F[j] = 0ull;
  for (unsigned long long j = i; j < na; j += st) A[j] = 0ull;
}

static __global__ void bitbfs_seed_kernel(bitbfs_params p, const unsigned* __restrict__ src, unsigned K,
                                   unsigned max_iter) {
  unsigned const k = threadIdx.x;
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
  }
  clear_ticket_board(p.ctl->board.w, k, blockDim.x);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void bitbfs_step_kernel(bitbfs_params p) {
  constexpr unsigned kW = BLOCK / 64;
  __shared__ unsigned heads[kW][64];
  __shared__ unsigned is_last;
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
  unsigned upd = 0;
  unsigned long long act_edges = 0, pairs = 0;
  unsigned const nw = gridDim.x * kW;
  for (unsigned k = blockIdx.x * kW + wib; k < p.nitems; k += nw) {
    unsigned const e0 = k * p.epi;
    unsigned const e1 = min(e0 + p.epi, p.m);
    unsigned vp = p.istart[(k * p.epi) >> 6];   // chunk starts, one per 64 edges
    unsigned cur = 0xffffffffu;          // carried destination (uniform)
    unsigned long long cacc[kBitWords] = {0ull, 0ull, 0ull, 0ull};
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
        }
      }
      // levels of the new rows, written by the whole wave
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
#pragma unroll
            for (unsigned q = 0; q < 4; ++q)
              if ((bits >> q) & 1u) row[q] = level;
          }
        }
      }
      // carry the final run into the next chunk
      if (cm) {
        unsigned const j = 63u - __builtin_clzll(cm);
        cur = readlane(key, j);
#pragma unroll
        for (unsigned w = 0; w < kBitWords; ++w) cacc[w] = readlane64(x[w], j);
      } else {
        cur = 0xffffffffu;
#pragma unroll
        for (unsigned w = 0; w < kBitWords; ++w) cacc[w] = 0ull;
      }
      vp += adv;
    }
  }

  // each updated row was counted by the lane that flushed it: sum the wave
  {
    unsigned const ut = wave_sum(upd);
    if (lane == 0 && ut) keep_live(atomicAdd(&p.stats[iter].updated, ut));
  }
  {
    unsigned long long const a = wave_sum64(act_edges);
    unsigned long long const q = wave_sum64(pairs);
    if (lane == 0) {
      if (a) keep_live(atomicAdd(&p.stats[iter].active_edges, a));
      if (q) keep_live(atomicAdd(&p.stats[iter].new_pairs, q));
    }
  }
  __syncthreads();
  if (t == 0) {
    __builtin_amdgcn_s_waitcnt(0);
    is_last = take_ticket(p.ctl->board.w);
  }
  __syncthreads();
  if (!is_last || t != 0) return;
  unsigned const u = __hip_atomic_load(&p.stats[iter].updated, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  p.ctl->iter = iter + 1;
  if (u == 0) p.ctl->done = 1;
  else if (iter + 1 >= p.ctl->max_iter) {
    p.ctl->done = 1;
    p.ctl->overflow = 1;
  }
}

// Writes every u32 level row of L from its byte row: one wave per vertex,
// lane l turning bytes 4l .. 4l+3 into one 16-B store.  A byte 0xFF means
// unreached, unless the run went deeper than 254 levels (`deep`) and the
// vertex's VIS bit is set: then the level was stored straight into L and is
// kept.
static __global__ void bitbfs_expand_kernel(unsigned* __restrict__ L,
                                            const unsigned* __restrict__ L8w,
                                            const unsigned long long* __restrict__ VIS,
                                            unsigned n, int deep) {
  unsigned const lane = lane_id();
  unsigned const wpb = blockDim.x / 64;
  for (unsigned v = blockIdx.x * wpb + (threadIdx.x >> 6); v < n; v += gridDim.x * wpb) {
    unsigned long long const at = static_cast<unsigned long long>(v) * (kBitSources / 4) + lane;
    unsigned const b = L8w[at];
    unsigned x[4];
#pragma unroll
    for (unsigned q = 0; q < 4; ++q) {
      unsigned const c = (b >> (8u * q)) & 0xffu;
      x[q] = c == 0xffu ? 0xFFFFFFFFu : c;
    }
    uint4* out = reinterpret_cast<uint4*>(L + static_cast<unsigned long long>(v) * kBitSources) + lane;
    if (deep && b != 0u) {
      unsigned long long const vw = VIS[static_cast<unsigned long long>(v) * kBitWords + (lane >> 4)];
      unsigned const vb = static_cast<unsigned>(vw >> ((4u * lane) & 63u)) & 0xfu;
      bool keep = false;
#pragma unroll
      for (unsigned q = 0; q < 4; ++q)
        keep = keep || (x[q] == 0xFFFFFFFFu && ((vb >> q) & 1u));
      if (keep) {
        uint4 const old = *out;
        unsigned const o[4] = {old.x, old.y, old.z, old.w};
#pragma unroll
        for (unsigned q = 0; q < 4; ++q)
          if (x[q] == 0xFFFFFFFFu && ((vb >> q) & 1u)) x[q] = o[q];
      }
    }
    *out = make_uint4(x[0], x[1], x[2], x[3]);
  }
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_BITBFS_KERNEL_HPP_
