// step_kernel.hpp -- the iteration kernels of the engine.
//
// One launch of step_kernel runs one iteration of the fixpoint.  The
// iteration index, the variant to run and the done flag live in a device
// control block written by the last workgroup of the previous launch, so a
// chunk of launches can be captured once in a hipGraph and replayed with no
// host round trip per iteration.
//
// Variants (all on a fixed grid, grid-stride over their work items):
//   CLASSIC  the reference algorithm (kites/device/kernels_single_device.cuh
//            :18-169): every vertex whose frontier bit is set is recomputed
//            from all its in-edges (init, then nbrComp + red per in-edge,
//            then upd), the wave expands the in-edges of its 64 vertices as
//            one flat list (merged expansion) and reduces each vertex's
//            partials with a segmented shuffle reduction.  In classic-only
//            mode it also marks the out-neighbours of updated vertices (or
//            of their groups when vertex-per-group > 1) in the next
//            frontier, as the reference does (:127-167).
//   PUSH_T/W/B  delta step over CSR out-edges from the sparse queue of the
//            vertices updated in the previous iteration; one lane / one
//            wave-merged group of 64 queue entries / one workgroup per
//            queue entry.  Destination values are lowered atomically.
//   PULL_T/W/B  delta step over CSC in-edges of every vertex, gathering only
//            from sources whose bit is set in the dense bitmap of vertices
//            updated in the previous iteration; one lane / one wave-merged
//            group of 64 vertices / one workgroup per vertex.  Each vertex is
//            written by its owner only and the updated-vertex bitmap word is
//            the owner wave's ballot (no atomics).
// Delta variants require a min-semilattice functor set (init copies, red is
// idempotent/commutative/associative, upd is "red moved the value"); the
// host probes the functors before allowing them.  Every variant writes the
// updated-vertex bitmap and queue for the next iteration and the per
// iteration statistics the schedulers read.
//
// User functors keep the reference signatures with volatile references
// (README.md:16-23).  Each lane owns LDS slots that the functors operate on,
// so no operand is forced into scratch memory.
#ifndef KITES_DETAIL_STEP_KERNEL_HPP_
#define KITES_DETAIL_STEP_KERNEL_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>

#include "../common.hpp"
#include "mlp_device.hpp"
#include "ticket.hpp"
#include "wave.hpp"

namespace kites {
namespace detail {

enum variant_id : unsigned {
  V_CLASSIC = 0,
  V_PUSH_T = 1,
  V_PUSH_W = 2,
  V_PUSH_B = 3,
  V_PUSH_LB = 4,
  V_PULL_T = 5,
  V_PULL_W = 6,
  V_PULL_B = 7,
  V_PULL_LB = 8,
  V_INIT_LB = 9,
  V_SPLIT = 10,   // near-far: move far-pile entries below the new threshold to the queue
  V_PERSIST = 11, // push run by the one-workgroup loop (persist_kernel.hpp)
  V_COUNT = 12
};

// Granularities of the delta variants: variant = base + granularity with
// base V_PUSH_T (push) or V_PULL_T (pull).
constexpr unsigned kGranularities = 4;

// Queue appends reserve queue slots and out-edge offsets with one 64-bit
// atomic: count in the high 32 bits, out-edge total in the low 32 bits, so
// queue order and edge-offset order agree.  A queue holds each vertex at
// most once (its bit in the next frontier decides), so the count is below
// 2^32 (n is a 32-bit index) and the out-edge total is at most m, which the
// 32-bit edge offsets bound by 2^32 - 1: neither field can carry into the
// other.
constexpr unsigned kQeShift = 32;
constexpr unsigned long long kQeMask = (1ull << kQeShift) - 1ull;
__host__ __device__ constexpr unsigned long long qe_pack(unsigned count, unsigned edges) {
  return (static_cast<unsigned long long>(count) << kQeShift) | edges;
}
__host__ __device__ constexpr unsigned qe_count(unsigned long long qe) {
  return static_cast<unsigned>(qe >> kQeShift);
}
__host__ __device__ constexpr unsigned qe_edges(unsigned long long qe) {
  return static_cast<unsigned>(qe & kQeMask);
}
// The largest queue (every vertex of a 2^32 - 1 vertex graph) with the
// largest edge total, reached in two partial appends, unpacks exactly.
static_assert(qe_count(qe_pack(0xFFFFFFFEu, 0x7FFFFFFFu) + qe_pack(1u, 0x80000000u)) ==
                  0xFFFFFFFFu,
              "queue count field");
static_assert(qe_edges(qe_pack(0xFFFFFFFEu, 0x7FFFFFFFu) + qe_pack(1u, 0x80000000u)) ==
                  0xFFFFFFFFu,
              "queue edge field");
static_assert(qe_count(qe_pack(1u << 26, 5u)) == (1u << 26), "counts above 2^26");

enum sched_mode : unsigned {
  SCHED_CLASSIC = 0,   // classic-only (generic functors)
  SCHED_PLAIN = 1,     // fixed push/pull threshold, fixed granularity
  SCHED_FORCED = 2,    // every delta iteration runs `forced_variant`
  SCHED_LEARNED = 3    // MLP picks the variant of every delta iteration
};

// Statistics of one iteration (written by the kernel of that iteration).
struct step_stats {
  unsigned count;               // vertices updated (= next queue length)
  unsigned variant;             // variant that ran
  unsigned max_outdeg;          // max out-degree over updated vertices
  unsigned in_count;            // queue length this iteration consumed
  unsigned long long sum_outdeg;  // out-edges of updated vertices
  unsigned long long work_edges;  // edges examined
  unsigned long long gathers;     // source values gathered
  unsigned long long active;      // vertices / queue entries processed
  unsigned long long qe;          // packed (count, out-edges) append counter
};
static_assert(sizeof(step_stats) == 56, "step_stats layout");

// Queue entries + out-edges below which a push iteration runs as fast on a
// grid of one workgroup per compute unit as on the full grid.
constexpr unsigned long long kSmallWork = 1ull << 18;
// At most this much work (entries + out-edges) the next chunk runs on the
// tiny grid (a quarter workgroup per compute unit): on 2048 x 2048 grid
// BFS (about 10^4 work per iteration) 64 workgroups take 10.3 us per
// iteration, 128 take 10.5, 256 take 11.1 and 16 take 12.1.
constexpr unsigned long long kTinyWork = 1ull << 16;

// A push iteration (thread, wave or edge-balanced) runs on
// ceil(work / kWorkPerGroup) workgroups, work = queue entries + out-edges,
// clamped to [one per compute unit, the whole grid].  Measured on R-MAT
// 21/22 SSSP and BFS: 0.5M work is fastest on 1-3 workgroups per CU, 1M on
// 2, 5M on 10, 20M and more on the whole grid (21 per CU); a larger grid
// than the work needs only adds fixed cost per workgroup.
constexpr unsigned long long kWorkPerGroup = 2048;

// The learned schedule may pull only if out-edges(frontier) * guard >= m:
// a pull reads every in-edge, a push only the frontier's out-edges (at
// about 2-3x the cost per edge: random reads and atomics), so a pull that
// walks whole in-lists wins only on frontiers whose out-edges are a large
// share of m (kPullGuard: 1/8); a settled run's pull stops at the first
// updated in-neighbour and reads far less (kPullGuardSettled: 1/32).
constexpr unsigned long long kPullGuard = 8;
constexpr unsigned long long kPullGuardSettled = 32;

// Per-group partial statistics of the running iteration: workgroup b adds
// into group b % kTicketGroups, so no address sees more than gridDim / 64
// atomics; the last workgroup folds the groups into the iteration's
// step_stats and clears them.  One 128-byte line per group: atomics on one
// line serialize.
struct alignas(128) group_stats {
  unsigned long long work, gathers, active;
  unsigned max_deg, pad;
  unsigned long long vmin;   // near-far: least value entering the next queue
};

struct step_ctl {
  unsigned iter;
  unsigned done;
  unsigned next_variant;
  unsigned max_iter;
  unsigned overflow;
  unsigned seed_count;   // queue length the first iteration consumes (seeded runs)
  unsigned small_next;   // host hint for the next chunk's grid: 0 full, 1 small, 2 tiny
  // The queue the next iteration consumes (written with `iter`, so that an
  // iteration reads its input from this one line): length and out-edges.
  unsigned q_count;
  unsigned long long q_edges;
  unsigned cache_key;    // learned: quantized features of the last inference
  unsigned cache_variant;  // and its decision
  // Near-far (options::delta > 0): relaxations to values below `thresh` go
  // to the queue, the others to the far pile (far_n entries in pile
  // far_sel; far_next counts the entries a split keeps in the other pile);
  // far_vmin is a lower bound of the far pile's values, exact after a split.
  unsigned long long thresh;
  unsigned long long thresh_prev;  // the threshold before the last split
  unsigned long long far_vmin;
  unsigned far_n;
  unsigned far_next;
  unsigned far_sel;
  ticket_board board;         // completion-ticket counters (not copied to the host)
};

struct sched_params {
  unsigned mode;
  unsigned fixed_gran;      // 0 thread, 1 wave, 2 workgroup, 3 edge-balanced
  unsigned forced_variant;
  float alpha;              // pull when out-edges(U) * alpha > m
  const float* mlp;         // learned model weights (device), mode LEARNED
  float margin;             // learned: keep the plain choice unless another variant's
                            // prediction is lower by more than this (log2 units)
  unsigned long long delta; // near-far bucket width (0 = off; unsigned min functor sets)
  // Settled values (unsigned min functor sets; 0 = off): a value below this
  // bound never changes once written, and every candidate a vertex is
  // offered in the iteration that first lowers it below the bound is the
  // same (BFS levels: the bound is the unreached value).  Pulls then skip
  // settled destinations and stop at the first updated in-neighbour.
  unsigned long long settled_below;
};

// Whether `v` is settled under the run's bound (never for signed or
// non-native-min value types).
template <class vT>
__device__ __forceinline__ bool is_settled(sched_params const& s, vT v) {
  if constexpr (std::is_unsigned<vT>::value)
    return s.settled_below != 0 && static_cast<unsigned long long>(v) < s.settled_below;
  else
    return false;
}

// Parts of the source-split pull: one per XCD of the MI355X.
constexpr unsigned kSplitParts = 8;

// Part of a source vertex in the split pull: a multiplicative hash of its
// 32-vertex line (128 B of 4-byte values), so that each part owns whole
// cache lines of V and the parts carry similar edge counts on graphs whose
// degrees fall with the vertex id (R-MAT).
__host__ __device__ __forceinline__ unsigned split_part(unsigned v) {
  return ((v >> 5) * 0x9E3779B1u) >> 29;
}

// Plain-data part of the kernel parameters.
template <class vT, class eT>
struct step_core {
  unsigned n;
  unsigned long long m;
  const unsigned* R;
  const unsigned* C;
  eT* E;
  const unsigned* oR;
  const unsigned* oC;
  eT* oE;
  const unsigned* nzi;  // vertices with in-edges, ascending
  const unsigned* nzR;  // nzR[i] = R[nzi[i]], nzR[nnz] = m
  unsigned nnz;
  const unsigned long long* outbits;  // bit v set iff v has out-edges
  const unsigned* gR;   // group out-lists (vertex-per-group > 1)
  const unsigned* gC;
  unsigned ngroups;     // groups that have an out-list (gR holds ngroups+1)
  vT* V;
  unsigned vpg_shift;
  unsigned vpg;
  // Rings indexed by iteration: 3 bitmaps of nwords words, 2 queues of n
  // entries, 2 offset arrays of n + 1 entries (out-edge offset of each queue
  // entry plus a sentinel).  One allocation each, so every pointer the
  // kernel derives stays a global-memory pointer.
  unsigned long long* bm;
  unsigned* q;
  unsigned* qe;
  unsigned nwords;
  // Near-far: two far piles of n entries and the far bitmap (nwords words).
  unsigned* fq;
  unsigned long long* fbm;
  step_stats* stats;
  step_ctl* ctl;
  unsigned* update_flag;
  sched_params sched;
  unsigned collect;     // collect diagnostic counters (work, gathers, active, max degree)
  group_stats* gst;     // kTicketGroups partial slots
  unsigned small_grid;  // least grid of a push iteration (0: push iterations use the whole grid)
  unsigned* board;      // completion-ticket counters (kTicketWords)
  unsigned max_outdeg;  // largest out-degree of the graph (0: unknown)
  unsigned persist_edges;  // persist_kernel: largest queue out-edge total it takes (0: off)
  // Source-split in-CSR for the edge-balanced pull (used when xsplit != 0):
  // the in-edges grouped by part of their source (kSplitParts parts, a
  // 32-vertex line of V belongs to one part), each part's edges ordered by
  // destination; xnzi/xnzR list each part's destination segments like
  // nzi/nzR, part b's segments being [xseg[b], xseg[b+1]).  Workgroup w
  // works on part w % kSplitParts: consecutive workgroups go to different
  // XCDs, so each XCD gathers source values from its own eighth of V, which
  // stays in its L2.
  const unsigned* xC;
  eT* xE;
  const unsigned* xnzi;
  const unsigned* xnzR;
  unsigned xseg[kSplitParts + 1];
  unsigned xsplit;
};

// Kernel parameters: the plain data plus the user's four functors (lambda
// types are not default-constructible, hence the constructor).
template <class vT, class eT, class FI, class FN, class FR, class FU>
struct step_params : step_core<vT, eT> {
  FI fi;
  FN fn;
  FR fr;
  FU fu;
  step_params(step_core<vT, eT> const& c, FI a, FN b, FR r, FU u)
      : step_core<vT, eT>(c), fi(a), fn(b), fr(r), fu(u) {}
};

constexpr int kWaves = COMPILE_TIME_DETERMINED_BLOCK_SIZE / WAVE_SIZE;
// Staged queue entries per wave (at least 64 * the items per lane of any
// caller of stage_queue).
constexpr unsigned kAcap = 256;
// Edges per lane of one chunk of the edge-balanced variants (lane l takes
// the IPL consecutive edges at l * IPL).  1 measures fastest on MI355X: at
// 7 waves per SIMD the other waves hide the load latency, and a smaller
// chunk wastes fewer lanes in the last, partial chunk of each 64-item
// window (R-MAT 21/22 SSSP and BFS: IPL 1 < 2 < 3 < 4 < 8 in run time).
constexpr unsigned kIPL = 1;
constexpr unsigned kChunk = WAVE_SIZE * kIPL;

template <class vT>
struct step_smem {
  vT acc[COMPILE_TIME_DETERMINED_BLOCK_SIZE];
  vT tmp[COMPILE_TIME_DETERMINED_BLOCK_SIZE];
  vT rhs[COMPILE_TIME_DETERMINED_BLOCK_SIZE];
  vT val[kWaves][WAVE_SIZE];
  vT slot[kWaves][WAVE_SIZE];
  unsigned excl[kWaves][WAVE_SIZE];
  unsigned base[kWaves][WAVE_SIZE];
  unsigned item[kWaves][WAVE_SIZE];
  unsigned win[kWaves][WAVE_SIZE + 8];
  unsigned hasf[kWaves][WAVE_SIZE];
  vT wpart[kWaves];
  unsigned whas[kWaves];
  unsigned long long ubits;
  int is_last;
  // Queue appends are staged per wave and written with one global atomic
  // per workgroup (or per full wave buffer): abuf_v/abuf_e hold the staged
  // vertices and their out-edge offsets relative to the wave's buffer,
  // ahdr the wave's staged count and out-edge total.  Near-far runs split
  // each wave's buffer: the queue stages in abuf_v/abuf_e[0, kAcap/2), the
  // far pile in abuf_v[kAcap/2, kAcap) with its count in ahdr[.][2].
  unsigned abuf_v[kWaves][kAcap];
  unsigned abuf_e[kWaves][kAcap];
  unsigned ahdr[kWaves][3];
  unsigned long long bbase;
  unsigned fbase;       // near-far: the workgroup's first far-pile slot
  unsigned long long bstat[3];
  unsigned bmax;
  unsigned long long bvmin;
  vT bval;              // PUSH_B: the queue entry's value, read once per workgroup
};

struct iter_ctx {
  unsigned iter;
  unsigned grid;      // workgroups that share this iteration's work (the first `grid`)
  const unsigned long long* in_bm;
  unsigned long long* out_bm;
  unsigned long long* clr_bm;
  const unsigned* in_q;
  unsigned* out_q;
  const unsigned* in_qe;
  unsigned* out_qe;
  unsigned in_count;
  unsigned long long in_edges;   // push: out-edges of the input queue (its sentinel offset)
  bool tiny;          // push: entries + out-edges <= kTinyWork (bits claimed without a read)
  step_stats* st;
  bool scatter;
  bool dense_in;      // the updated set's out-edges cover most edges: pull gathers
                      // every in-edge without testing the bitmap
  bool split;         // the edge-balanced pull runs over the source-split in-CSR
  bool no_deg;        // learned: the next decision is surely guarded, skip the degree statistics
  unsigned variant;   // the variant this iteration runs
  bool nf;            // near-far: relaxations to values >= T go to the far pile
  unsigned long long T;
  unsigned long long Told;   // split: the threshold before this split
  unsigned* far_q;           // the far pile appends go to (a split: the other pile)
  const unsigned* far_src;   // split: the pile it reads (the current one)
  unsigned long long* far_bm;
  unsigned* far_n;           // the current pile's append counter
  unsigned* abuf_v;   // the calling wave's staging area (LDS)
  unsigned* abuf_e;
  unsigned* ahdr;
  unsigned acap;      // queue entries a wave stages before it flushes
  unsigned* fbuf;     // near-far: the wave's far-pile staging area (LDS)
};

struct lane_acc {
  unsigned long long work = 0, gathers = 0, active = 0;
  unsigned max_deg = 0;
  unsigned long long vmin = ~0ull;   // near-far: least value appended by this lane
};

// Near-far push (unsigned min functor sets, options::delta > 0).  A
// relaxation that lowers V[v] below the threshold T queues v for the next
// iteration (the near queue); one that leaves V[v] at or above T puts v on
// the far pile instead (once: the far bitmap).  Far entries are not
// expanded, so their out-edges cost nothing until the near queue runs dry;
// then a split iteration (V_SPLIT) raises T to the least far value + delta
// and moves the far entries below it to the queue.
template <class vT>
__device__ __forceinline__ bool is_near(iter_ctx const& c, vT v) {
  return !c.nf || static_cast<unsigned long long>(v) < c.T;
}

// Far-pile bookkeeping: the least value of a far relaxation.
template <class vT>
__device__ __forceinline__ void note_far(lane_acc& la, vT v) {
  unsigned long long const x = static_cast<unsigned long long>(v);
  la.vmin = x < la.vmin ? x : la.vmin;
}

template <class T>
__device__ __forceinline__ volatile T& vs(T& x) {
  return *reinterpret_cast<volatile T*>(&x);
}

// Keeps a loaded value alive without using it (forces the load).
template <class T>
__device__ __forceinline__ void keep_live(T const& x) {
  static_assert(sizeof(T) % 4 == 0, "4-byte granular value");
  unsigned w[sizeof(T) / 4];
  __builtin_memcpy(w, &x, sizeof(T));
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) asm volatile("" ::"v"(w[i]));
}

__device__ __forceinline__ bool test_bit(const unsigned long long* bm, unsigned i) {
  return (bm[i >> 6] >> (i & 63)) & 1ull;
}

// Sets bit i; true iff this call changed it from 0 to 1.  Bits only go
// from 0 to 1 while a bitmap is written, so a plain (possibly stale) read
// that already shows the bit proves it set; otherwise the returning
// atomicOr decides.
__device__ __forceinline__ bool set_bit_once(unsigned long long* bm, unsigned i) {
  unsigned long long const bit = 1ull << (i & 63);
  unsigned long long* w = bm + (i >> 6);
  if (*reinterpret_cast<volatile unsigned long long*>(w) & bit) return false;
  unsigned long long old = atomicOr(w, bit);
  return !(old & bit);
}

// Sets bit i with one returning atomic (no read first: one round trip on
// iterations too small for the saved atomics to matter); true iff this call
// changed it from 0 to 1.
__device__ __forceinline__ bool set_bit_now(unsigned long long* bm, unsigned i) {
  unsigned long long const bit = 1ull << (i & 63);
  return !(atomicOr(bm + (i >> 6), bit) & bit);
}

__device__ __forceinline__ void set_bit(unsigned long long* bm, unsigned i) {
  unsigned long long const bit = 1ull << (i & 63);
  unsigned long long* w = bm + (i >> 6);
  if (!(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit)) atomicOr(w, bit);
}

// True if v has out-edges.  Only such vertices enter the next frontier: a
// vertex without out-edges is never a source, so its change reaches nobody.
template <class P>
__device__ __forceinline__ bool has_out(P& p, unsigned v) {
  return (p.outbits[v >> 6] >> (v & 63)) & 1ull;
}

// Marks the vertices that have out-edges (one thread per 64-vertex word).
static __global__ void outbits_kernel(const unsigned* __restrict__ oR, unsigned n,
                                      unsigned long long* __restrict__ bits, unsigned nwords) {
  unsigned const w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  unsigned long long word = 0;
  for (unsigned j = 0; j < 64; ++j) {
    unsigned const v = w * 64 + j;
    if (v < n && oR[v + 1] != oR[v]) word |= 1ull << j;
  }
  bits[w] = word;
}

// Largest out-degree: *out = max over v < n of oR[v+1] - oR[v] (*out starts at 0).
static __global__ void max_outdeg_kernel(const unsigned* __restrict__ oR, unsigned n,
                                         unsigned* __restrict__ out) {
  unsigned const stride = gridDim.x * blockDim.x;
  unsigned md = 0;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    unsigned const d = oR[v + 1] - oR[v];
    md = d > md ? d : md;
  }
  md = wave_max(md);
  if ((threadIdx.x & 63) == 0 && md) atomicMax(out, md);
}

// ---- native min for the min-semilattice fast path --------------------------
template <class vT>
struct min_ops {
  static constexpr bool ok = std::is_integral<vT>::value && (sizeof(vT) == 4 || sizeof(vT) == 8);
};

template <class vT>
__device__ __forceinline__ vT atomic_min_any(vT* p, vT v) {
  if constexpr (sizeof(vT) == 4) {
    if constexpr (std::is_signed<vT>::value)
      return static_cast<vT>(atomicMin(reinterpret_cast<int*>(p), static_cast<int>(v)));
    else
      return static_cast<vT>(atomicMin(reinterpret_cast<unsigned*>(p), static_cast<unsigned>(v)));
  } else {
    if constexpr (std::is_signed<vT>::value)
      return static_cast<vT>(atomicMin(reinterpret_cast<long long*>(p), static_cast<long long>(v)));
    else
      return static_cast<vT>(atomicMin(reinterpret_cast<unsigned long long*>(p),
                                       static_cast<unsigned long long>(v)));
  }
}

template <class vT>
__device__ __forceinline__ vT atomic_cas_global(vT* p, vT cmp, vT val) {
  static_assert(sizeof(vT) == 4 || sizeof(vT) == 8, "CAS needs a 4- or 8-byte value");
  if constexpr (sizeof(vT) == 4) {
    unsigned c, v2, r;
    __builtin_memcpy(&c, &cmp, 4);
    __builtin_memcpy(&v2, &val, 4);
    r = atomicCAS(reinterpret_cast<unsigned*>(p), c, v2);
    vT out;
    __builtin_memcpy(&out, &r, 4);
    return out;
  } else {
    unsigned long long c, v2, r;
    __builtin_memcpy(&c, &cmp, 8);
    __builtin_memcpy(&v2, &val, 8);
    r = atomicCAS(reinterpret_cast<unsigned long long*>(p), c, v2);
    vT out;
    __builtin_memcpy(&out, &r, 8);
    return out;
  }
}

// ---- helpers shared by the variants ----------------------------------------

// Writes the calling wave's staged queue entries to the next queue with one
// global atomic (the packed counter reserves queue slots and out-edge
// offsets together).  All 64 lanes active.
__device__ __forceinline__ void wave_flush(iter_ctx& c) {
  unsigned const lane = lane_id();
  wave_lds_sync();
  unsigned const cnt = c.ahdr[0], edg = c.ahdr[1];
  if (!cnt) return;
  unsigned long long old = 0;
  if (lane == 0)
    old = atomicAdd(&c.st->qe, qe_pack(cnt, edg));
  unsigned const olo = readlane(static_cast<unsigned>(old), 0);
  unsigned const ohi = readlane(static_cast<unsigned>(old >> 32), 0);
  unsigned long long const o64 = (static_cast<unsigned long long>(ohi) << 32) | olo;
  unsigned const bc = qe_count(o64);
  unsigned const be = qe_edges(o64);
  for (unsigned i = lane; i < cnt; i += WAVE_SIZE) {
    c.out_q[bc + i] = c.abuf_v[i];
    c.out_qe[bc + i] = be + c.abuf_e[i];
  }
  wave_lds_sync();
  if (lane == 0) {
    c.ahdr[0] = 0;
    c.ahdr[1] = 0;
  }
  wave_lds_sync();
}

// Stages the lane's vertices v[i] with want[i] for the next queue, with
// their out-degrees.  All 64 lanes active.
template <class P, unsigned N>
__device__ __forceinline__ void stage_queue(P& p, iter_ctx& c, lane_acc& la, const bool (&want)[N],
                                            const unsigned (&v)[N]) {
  // One call stages up to 64 * N entries after at most one flush.
  static_assert(WAVE_SIZE * N <= kAcap, "stage_queue: a wave's entries exceed the staging buffer");
  unsigned cnt = 0, dsum = 0;
  unsigned deg[N];
#pragma unroll
  for (unsigned i = 0; i < N; ++i) {
    deg[i] = 0;
    if (want[i]) {
      deg[i] = p.oR[v[i] + 1] - p.oR[v[i]];
      ++cnt;
      dsum += deg[i];
    }
  }
  if (!ballot64(cnt != 0)) return;
  unsigned const ic = wave_incl_scan(cnt);
  unsigned const idg = wave_incl_scan(dsum);
  unsigned const tc = readlane(ic, 63), td = readlane(idg, 63);
  wave_lds_sync();
  if (c.ahdr[0] + tc > c.acap) wave_flush(c);
  unsigned const hc = c.ahdr[0], he = c.ahdr[1];
  unsigned pos = hc + ic - cnt;
  unsigned eo = he + idg - dsum;
#pragma unroll
  for (unsigned i = 0; i < N; ++i) {
    if (want[i]) {
      c.abuf_v[pos] = v[i];
      c.abuf_e[pos] = eo;
      ++pos;
      eo += deg[i];
      la.max_deg = deg[i] > la.max_deg ? deg[i] : la.max_deg;
    }
  }
  wave_lds_sync();
  if (lane_id() == 0) {
    c.ahdr[0] = hc + tc;
    c.ahdr[1] = he + td;
  }
  wave_lds_sync();
}

template <class P>
__device__ __forceinline__ void append_queue(P& p, iter_ctx& c, lane_acc& la, bool want,
                                             unsigned v) {
  bool const w[1] = {want};
  unsigned const vv[1] = {v};
  stage_queue(p, c, la, w, vv);
}

// Writes the calling wave's staged far-pile entries with one global atomic
// on the pile's counter.  All 64 lanes active.
__device__ __forceinline__ void far_wave_flush(iter_ctx& c) {
  unsigned const lane = lane_id();
  wave_lds_sync();
  unsigned const cnt = c.ahdr[2];
  if (!cnt) return;
  unsigned base = 0;
  if (lane == 0) base = atomicAdd(c.far_n, cnt);
  base = readlane(base, 0);
  for (unsigned i = lane; i < cnt; i += WAVE_SIZE) c.far_q[base + i] = c.fbuf[i];
  wave_lds_sync();
  if (lane == 0) c.ahdr[2] = 0;
  wave_lds_sync();
}

// Stages the lanes' vertices with `want` set for the far pile; the pile's
// counter sees one atomic per full wave buffer or per workgroup
// (block_flush), as the queue's does.  All 64 lanes active.
__device__ __forceinline__ void far_append(iter_ctx& c, bool want, unsigned v) {
  unsigned long long const m = ballot64(want);
  if (!m) return;
  unsigned const cnt = static_cast<unsigned>(__builtin_popcountll(m));
  wave_lds_sync();
  if (c.ahdr[2] + cnt > kAcap / 2) far_wave_flush(c);
  unsigned const h = c.ahdr[2];
  if (want) c.fbuf[h + rank_in(m)] = v;
  wave_lds_sync();
  if (lane_id() == 0) c.ahdr[2] = h + cnt;
  wave_lds_sync();
}

// After a push relaxation that may have lowered V[v] (v has out-edges) to
// `cand`: a near value claims v's bit in the next frontier (true: append v
// to the queue); a far value claims v's far bit instead (`far`: append v to
// the far pile).
template <class vT>
__device__ __forceinline__ bool claim_next(iter_ctx& c, lane_acc& la, unsigned v, vT cand,
                                           bool& far) {
  if (is_near(c, cand)) return c.tiny ? set_bit_now(c.out_bm, v) : set_bit_once(c.out_bm, v);
  note_far(la, cand);
  far = set_bit_once(c.far_bm, v);
  return false;
}

// Publishes the updated vertices of one owner wave's 64-vertex word: the
// word of the next frontier bitmap is the wave's ballot (no atomics), and
// the vertices are staged for the next queue.  All 64 lanes active.
template <class P>
__device__ __forceinline__ void publish_word(P& p, iter_ctx& c, lane_acc& la, bool upd,
                                             unsigned v, unsigned word) {
  bool const keep = upd && has_out(p, v);
  unsigned long long const km = ballot64(keep);
  if (!km) return;
  if (lane_id() == 0) c.out_bm[word] = km;
  append_queue(p, c, la, keep, v);
}

// Writes every wave's staged entries of the workgroup with one global
// atomic.  Whole workgroup, uniform control flow.
// With gst, also adds the workgroup's statistics (flush_stats) into its
// group's slot; those atomics are issued first so that their round trips
// overlap the queue counter's.
template <class vT>
__device__ __forceinline__ void block_flush(iter_ctx& c, step_smem<vT>& sm, group_stats* gst) {
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tc = 0, te = 0;
    for (int w = 0; w < kWaves; ++w) {
      tc += sm.ahdr[w][0];
      te += sm.ahdr[w][1];
    }
    // Returning atomics: the ticket that follows waits for their results.
    unsigned long long r0 = 0, r1 = 0, r2 = 0, r4 = 0;
    unsigned r3 = 0;
    if (gst) {
      unsigned const ng = gridDim.x < kTicketGroups ? gridDim.x : kTicketGroups;
      group_stats* g = gst + blockIdx.x % ng;
      if (sm.bstat[0]) r0 = atomicAdd(&g->work, sm.bstat[0]);
      if (sm.bstat[1]) r1 = atomicAdd(&g->gathers, sm.bstat[1]);
      if (sm.bstat[2]) r2 = atomicAdd(&g->active, sm.bstat[2]);
      if (sm.bmax) r3 = atomicMax(&g->max_deg, sm.bmax);
      if (sm.bvmin != ~0ull) r4 = atomicMin(&g->vmin, sm.bvmin);
    }
    unsigned long long old = 0;
    if (tc) old = atomicAdd(&c.st->qe, qe_pack(tc, te));
    sm.bbase = old;
    if (c.nf) {
      unsigned tf = 0;
      for (int w = 0; w < kWaves; ++w) tf += sm.ahdr[w][2];
      sm.fbase = tf ? atomicAdd(c.far_n, tf) : 0u;
    }
    keep_live(r0);
    keep_live(r1);
    keep_live(r2);
    keep_live(r3);
    keep_live(r4);
  }
  __syncthreads();
  unsigned const wib = threadIdx.x >> 6, lane = lane_id();
  unsigned long long const o64 = sm.bbase;
  unsigned bc = qe_count(o64);
  unsigned be = qe_edges(o64);
  for (unsigned w = 0; w < wib; ++w) {
    bc += sm.ahdr[w][0];
    be += sm.ahdr[w][1];
  }
  unsigned const cnt = sm.ahdr[wib][0];
  for (unsigned i = lane; i < cnt; i += WAVE_SIZE) {
    c.out_q[bc + i] = sm.abuf_v[wib][i];
    c.out_qe[bc + i] = be + sm.abuf_e[wib][i];
  }
  if (c.nf) {
    unsigned fb = sm.fbase;
    for (unsigned w = 0; w < wib; ++w) fb += sm.ahdr[w][2];
    unsigned const fc = sm.ahdr[wib][2];
    for (unsigned i = lane; i < fc; i += WAVE_SIZE) c.far_q[fb + i] = c.fbuf[i];
  }
}

template <bool MIN_FAST, class P, class vT>
__device__ __forceinline__ bool relax(P& p, unsigned v, vT& cand, vT& work);

// Push-side relax of V[v] with the partial value in `cand`.  With native
// min the atomic does not return: the result is "v may have moved" (the
// candidate beat a plain read of V[v]), which at worst enqueues v once
// more than needed -- an extra, harmless relaxation next iteration -- and
// saves the atomic's round trip.
template <bool MIN_FAST, class P, class vT>
__device__ __forceinline__ bool relax_push(P& p, unsigned v, vT& cand, vT& work) {
  if constexpr (MIN_FAST) {
    vT const cv = vs(cand);
    if (!(cv < p.V[v])) return false;
    atomic_min_any(p.V + v, cv);
    return true;
  } else {
    return relax<false>(p, v, cand, work);
  }
}

// Relaxes V[v] with the partial value in `cand` (LDS slot); true if V[v]
// moved.  Delta variants only.
template <bool MIN_FAST, class P, class vT>
__device__ __forceinline__ bool relax(P& p, unsigned v, vT& cand, vT& work) {
  if constexpr (MIN_FAST) {
    vT const cv = vs(cand);
    if (!(cv < p.V[v])) return false;
    vT const old = atomic_min_any(p.V + v, cv);
    return cv < old;
  } else {
    vT old = vs(p.V[v]);
    for (;;) {
      vs(work) = old;
      p.fr(vs(work), vs(cand));
      if (!p.fu(vs(work), old)) return false;
      vT const nv = vs(work);
      vT const prev = atomic_cas_global(p.V + v, old, nv);
      if (prev == old) return true;
      old = prev;
    }
  }
}

// Segmented reduction of the chunk's partials towards each segment head.
// key: owner lane (contiguous runs), has: lane holds a partial in tmp[t].
// Returns has of the calling lane after the reduction.
template <class P, class vT>
__device__ __forceinline__ bool seg_reduce(P& p, step_smem<vT>& sm, unsigned key, bool has) {
  unsigned const lane = lane_id(), t = threadIdx.x;
#pragma unroll 1
  for (unsigned s = 1; s < 64; s <<= 1) {
    vT const x = sm.tmp[t];
    vT const y = shfl_down_any(x, s);
    unsigned const ky = __shfl_down(key, s, 64);
    unsigned const hy = __shfl_down((unsigned)has, s, 64);
    bool const same = (lane + s < 64) && (ky == key);
    if (!ballot64(same)) break;
    if (same && hy) {
      if (has) {
        sm.rhs[t] = y;
        p.fr(vs(sm.tmp[t]), vs(sm.rhs[t]));
      } else {
        sm.tmp[t] = y;
        has = true;
      }
    }
  }
  return has;
}

// Adds a lane's statistics into the workgroup's LDS counters (one LDS
// atomic per wave per counter); block_flush writes them out.
__device__ __forceinline__ unsigned long long wave_min64(unsigned long long x) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    unsigned long long const y = __shfl_xor(x, s, 64);
    x = y < x ? y : x;
  }
  return x;
}

template <class vT>
__device__ __forceinline__ void flush_stats(step_smem<vT>& sm, lane_acc& la, bool full, bool nf) {
  if (nf) {
    unsigned long long const vm = wave_min64(la.vmin);
    if (lane_id() == 0 && vm != ~0ull) atomicMin(&sm.bvmin, vm);
  }
  unsigned const md = wave_max(la.max_deg);
  if (full) {
    unsigned long long const wk = wave_sum64(la.work);
    unsigned long long const ga = wave_sum64(la.gathers);
    unsigned long long const ac = wave_sum64(la.active);
    if (lane_id() == 0) {
      if (wk) atomicAdd(&sm.bstat[0], wk);
      if (ga) atomicAdd(&sm.bstat[1], ga);
      if (ac) atomicAdd(&sm.bstat[2], ac);
    }
  }
  if (lane_id() == 0 && md) atomicMax(&sm.bmax, md);
}

// Last workgroup, first wave: folds the group partials into `st` (all 64
// lanes active) and clears them for the next launch; returns max degree.
__device__ __forceinline__ unsigned fold_group_stats(group_stats* gst, step_stats* st, bool full,
                                                     bool want_vmin, unsigned long long& vmin) {
  unsigned const lane = lane_id();
  unsigned const ng = gridDim.x < kTicketGroups ? gridDim.x : kTicketGroups;
  unsigned md = 0;
  unsigned long long vm = ~0ull;
  if (lane < ng) {
    group_stats* g = gst + lane;
    md = __hip_atomic_load(&g->max_deg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    g->max_deg = 0;
    if (want_vmin) {
      vm = __hip_atomic_load(&g->vmin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      g->vmin = ~0ull;
    }
  }
  if (want_vmin) vmin = wave_min64(vm);
  md = wave_max(md);
  if (lane == 0) st->max_outdeg = md;
  if (full) {
    unsigned long long wk = 0, ga = 0, ac = 0;
    if (lane < ng) {
      group_stats* g = gst + lane;
      wk = __hip_atomic_load(&g->work, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ga = __hip_atomic_load(&g->gathers, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ac = __hip_atomic_load(&g->active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      g->work = 0;
      g->gathers = 0;
      g->active = 0;
    }
    wk = wave_sum64(wk);
    ga = wave_sum64(ga);
    ac = wave_sum64(ac);
    if (lane == 0) {
      st->work_edges = wk;
      st->gathers = ga;
      st->active = ac;
    }
  }
  return md;
}

// ---- CLASSIC: reference algorithm, wave-merged expansion --------------------
template <bool MIN_FAST, class P, class vT>
__device__ void run_classic(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  unsigned const lane = lane_id(), wib = threadIdx.x >> 6, t = threadIdx.x;
  unsigned const ntasks = (p.n + 63) >> 6;
  unsigned const stride = c.grid * kWaves;
  unsigned* const excl = sm.excl[wib];
  unsigned* const base = sm.base[wib];
  for (unsigned task = blockIdx.x * kWaves + wib; task < ntasks; task += stride) {
    unsigned const v = (task << 6) + lane;
    bool const inr = v < p.n;
    bool active;
    if (p.vpg_shift == 0) {
      active = inr && ((c.in_bm[task] >> lane) & 1ull);
    } else {
      active = inr && test_bit(c.in_bm, v >> p.vpg_shift);
    }
    if (!ballot64(active)) continue;
    vT old{};
    unsigned r0 = 0, deg = 0;
    if (active) {
      old = p.V[v];
      vT g = old;
      p.fi(vs(sm.acc[t]), g);
      r0 = p.R[v];
      deg = p.R[v + 1] - r0;
    }
    la.active += active;
    la.work += deg;
    la.gathers += deg;
    unsigned const incl = wave_incl_scan(deg);
    unsigned const total = readlane(incl, 63);
    excl[lane] = incl - deg;
    base[lane] = r0;
    wave_lds_sync();
    for (unsigned c0 = 0; c0 < total; c0 += 64) {
      unsigned const e = c0 + lane;
      bool const valid = e < total;
      unsigned own = 64 + lane;
      if (valid) {
        own = owner_search(excl, e);
        unsigned const idx = base[own] + (e - excl[own]);
        vT const sv = p.V[p.C[idx]];
        p.fn(vs(sm.tmp[t]), sv, p.E + idx);
      }
      if constexpr (MIN_FAST) {
        if (valid) atomic_min_any(&sm.acc[(wib << 6) + own], sm.tmp[t]);
      } else {
        bool const has = seg_reduce(p, sm, own, valid);
        unsigned const kprev = __shfl_up(own, 1, 64);
        bool const head = valid && has && (lane == 0 || kprev != own);
        if (head) p.fr(vs(sm.acc[(wib << 6) + own]), vs(sm.tmp[t]));
      }
      wave_lds_sync();
    }
    bool upd = false;
    if (active) {
      upd = p.fu(vs(sm.acc[t]), old);
      if (upd) p.V[v] = sm.acc[t];
    }
    unsigned long long const umask = ballot64(upd);
    if (!umask) continue;
    if (c.scatter) {
      // Mark out-neighbours (vertex level) or out-groups (group level) in the
      // next frontier (kernels_single_device.cuh:127-167, :288-324).
      bool resp = upd;
      unsigned lo = 0, dg = 0;
      if (p.vpg_shift == 0) {
        if (upd) {
          lo = p.oR[v];
          dg = p.oR[v + 1] - lo;
        }
      } else {
        unsigned const g = v >> p.vpg_shift;
        unsigned long long gm;
        if (p.vpg >= 64) {
          gm = ~0ull;
        } else {
          unsigned long long const ones = (1ull << p.vpg) - 1ull;
          gm = ones << (lane & ~(p.vpg - 1u));
        }
        unsigned long long const below = (1ull << lane) - 1ull;
        resp = upd && !(umask & gm & below) && g < p.ngroups;
        if (resp) {
          lo = p.gR[g];
          dg = p.gR[g + 1] - lo;
        }
      }
      unsigned const inc2 = wave_incl_scan(dg);
      unsigned const tot2 = readlane(inc2, 63);
      wave_lds_sync();
      excl[lane] = inc2 - dg;
      base[lane] = lo;
      wave_lds_sync();
      const unsigned* const list = p.vpg_shift == 0 ? p.oC : p.gC;
      for (unsigned c0 = 0; c0 < tot2; c0 += 64) {
        unsigned const e = c0 + lane;
        if (e < tot2) {
          unsigned const own = owner_search(excl, e);
          set_bit(c.out_bm, list[base[own] + (e - excl[own])]);
        }
      }
      wave_lds_sync();
      if (lane == 0) {
        unsigned const r = atomicAdd(&c.st->count, (unsigned)__builtin_popcountll(umask));
        keep_live(r);
        *p.update_flag = 1u;
      }
    } else {
      publish_word(p, c, la, upd, v, task);
    }
  }
}

// ---- PUSH: CSR out-edges from the sparse queue ------------------------------
template <bool MIN_FAST, class P, class vT>
__device__ void run_push_thread(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  unsigned const lane = lane_id(), t = threadIdx.x;
  unsigned const nthreads = c.grid * blockDim.x;
  for (unsigned kb = blockIdx.x * blockDim.x + (t & ~63u); kb < c.in_count; kb += nthreads) {
    unsigned const k = kb + lane;
    unsigned e0 = 0, deg = 0, u = 0;
    vT val{};
    if (k < c.in_count) {
      u = c.in_q[k];
      val = p.V[u];
      e0 = p.oR[u];
      deg = p.oR[u + 1] - e0;
      la.active += 1;
      la.work += deg;
    }
    unsigned const maxd = wave_max(deg);
    for (unsigned i = 0; i < maxd; ++i) {
      bool newly = false, far = false;
      unsigned dst = 0;
      if (i < deg) {
        unsigned const e = e0 + i;
        dst = p.oC[e];
        p.fn(vs(sm.tmp[t]), val, p.oE + e);
        if (relax_push<MIN_FAST>(p, dst, sm.tmp[t], sm.rhs[t]) && has_out(p, dst))
          newly = claim_next(c, la, dst, vs(sm.tmp[t]), far);
      }
      append_queue(p, c, la, newly, dst);
      if (c.nf) far_append(c, far, dst);
    }
  }
}

template <bool MIN_FAST, class P, class vT>
__device__ void run_push_wave(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  unsigned const lane = lane_id(), wib = threadIdx.x >> 6, t = threadIdx.x;
  unsigned const ntasks = (c.in_count + 63) >> 6;
  unsigned const stride = c.grid * kWaves;
  unsigned* const excl = sm.excl[wib];
  unsigned* const base = sm.base[wib];
  vT* const vals = sm.val[wib];
  for (unsigned task = blockIdx.x * kWaves + wib; task < ntasks; task += stride) {
    unsigned const k = (task << 6) + lane;
    unsigned r0 = 0, deg = 0, u = 0;
    vT val{};
    if (k < c.in_count) {
      u = c.in_q[k];
      val = p.V[u];
      r0 = p.oR[u];
      deg = p.oR[u + 1] - r0;
      la.active += 1;
    }
    la.work += deg;
    unsigned const incl = wave_incl_scan(deg);
    unsigned const total = readlane(incl, 63);
    wave_lds_sync();
    excl[lane] = incl - deg;
    base[lane] = r0;
    vals[lane] = val;
    wave_lds_sync();
    for (unsigned c0 = 0; c0 < total; c0 += 64) {
      unsigned const e = c0 + lane;
      bool newly = false, far = false;
      unsigned dst = 0;
      if (e < total) {
        unsigned const own = owner_search(excl, e);
        unsigned const idx = base[own] + (e - excl[own]);
        dst = p.oC[idx];
        p.fn(vs(sm.tmp[t]), vals[own], p.oE + idx);
        if (relax_push<MIN_FAST>(p, dst, sm.tmp[t], sm.rhs[t]) && has_out(p, dst))
          newly = claim_next(c, la, dst, vs(sm.tmp[t]), far);
      }
      append_queue(p, c, la, newly, dst);
      if (c.nf) far_append(c, far, dst);
    }
  }
}

template <bool MIN_FAST, class P, class vT>
__device__ void run_push_block(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  unsigned const t = threadIdx.x;
  for (unsigned k = blockIdx.x; k < c.in_count; k += c.grid) {
    unsigned const u = c.in_q[k];
    // One read of V[u] for the whole workgroup, so that every thread relaxes
    // with the same value.
    if (t == 0) sm.bval = p.V[u];
    __syncthreads();
    vT const val = sm.bval;
    __syncthreads();
    unsigned const e0 = p.oR[u], e1 = p.oR[u + 1];
    if (t == 0) la.active += 1;
    for (unsigned eb = e0; eb < e1; eb += blockDim.x) {
      unsigned const e = eb + t;
      bool newly = false, far = false;
      unsigned dst = 0;
      if (e < e1) {
        la.work += 1;
        dst = p.oC[e];
        p.fn(vs(sm.tmp[t]), val, p.oE + e);
        if (relax_push<MIN_FAST>(p, dst, sm.tmp[t], sm.rhs[t]) && has_out(p, dst))
          newly = claim_next(c, la, dst, vs(sm.tmp[t]), far);
      }
      append_queue(p, c, la, newly, dst);
      if (c.nf) far_append(c, far, dst);
    }
  }
}

// ---- SPLIT (near-far): the far pile after the queue ran dry ----------------
// Raises nothing itself (the previous iteration's last workgroup moved the
// threshold from c.Told to c.T).  Every far entry u: V[u] < Told means its
// last decrease went through the queue, so it was expanded -- dropped;
// Told <= V[u] < T -- moved to the queue (it is expanded next iteration);
// otherwise kept, in the other pile.  Entries leaving the pile clear their
// far bit.  No edge is expanded.
template <class P, class vT>
__device__ void run_split(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>&) {
  unsigned const lane = lane_id();
  unsigned const nthreads = c.grid * blockDim.x;
  for (unsigned kb = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); kb < c.in_count;
       kb += nthreads) {
    unsigned const k = kb + lane;
    bool to_near = false, keep = false;
    unsigned u = 0;
    if (k < c.in_count) {
      u = c.far_src[k];
      unsigned long long const x = static_cast<unsigned long long>(p.V[u]);
      keep = x >= c.T;
      if (keep) {
        note_far(la, x);
      } else {
        atomicAnd(c.far_bm + (u >> 6), ~(1ull << (u & 63)));
        to_near = x >= c.Told && set_bit_once(c.out_bm, u);
      }
      la.active += 1;
    }
    append_queue(p, c, la, to_near, u);
    far_append(c, keep, u);   // into the other pile (see step_kernel)
  }
}

// ---- PULL: CSC in-edges of every vertex, dense bitmap of updated sources ---
template <bool MIN_FAST, class P, class vT>
__device__ void run_pull_thread(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  unsigned const lane = lane_id(), wib = threadIdx.x >> 6, t = threadIdx.x;
  unsigned const ntasks = (p.n + 63) >> 6;
  unsigned const stride = c.grid * kWaves;
  for (unsigned task = blockIdx.x * kWaves + wib; task < ntasks; task += stride) {
    unsigned const v = (task << 6) + lane;
    bool upd = false;
    if (v < p.n) {
      vT const old = p.V[v];
      unsigned const e0 = p.R[v], e1 = p.R[v + 1];
      la.active += 1;
      if constexpr (MIN_FAST) {
        vT acc = old;
        if (is_settled(p.sched, old)) {
          // settled: nothing can lower it
        } else if (p.sched.settled_below) {
          // every candidate is the same: the first updated in-neighbour's
          unsigned e = e0;
          for (; e < e1; ++e) {
            unsigned const s = p.C[e];
            if (test_bit(c.in_bm, s)) {
              la.gathers += 1;
              p.fn(vs(sm.tmp[t]), p.V[s], p.E + e);
              vT const cv = sm.tmp[t];
              acc = cv < acc ? cv : acc;
              ++e;
              break;
            }
          }
          la.work += e - e0;
        } else {
          la.work += e1 - e0;
          for (unsigned e = e0; e < e1; ++e) {
            unsigned const s = p.C[e];
            if (test_bit(c.in_bm, s)) {
              la.gathers += 1;
              p.fn(vs(sm.tmp[t]), p.V[s], p.E + e);
              vT const cv = sm.tmp[t];
              acc = cv < acc ? cv : acc;
            }
          }
        }
        upd = acc < old;
        if (upd) p.V[v] = acc;
      } else {
        la.work += e1 - e0;
        vT g = old;
        p.fi(vs(sm.acc[t]), g);
        for (unsigned e = e0; e < e1; ++e) {
          unsigned const s = p.C[e];
          if (test_bit(c.in_bm, s)) {
            la.gathers += 1;
            p.fn(vs(sm.tmp[t]), p.V[s], p.E + e);
            p.fr(vs(sm.acc[t]), vs(sm.tmp[t]));
          }
        }
        upd = p.fu(vs(sm.acc[t]), old);
        if (upd) p.V[v] = sm.acc[t];
      }
    }
    publish_word(p, c, la, upd, v, task);
  }
}

template <bool MIN_FAST, class P, class vT>
__device__ void run_pull_wave(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  unsigned const lane = lane_id(), wib = threadIdx.x >> 6, t = threadIdx.x;
  unsigned const ntasks = (p.n + 63) >> 6;
  unsigned const stride = c.grid * kWaves;
  unsigned* const excl = sm.excl[wib];
  unsigned* const base = sm.base[wib];
  for (unsigned task = blockIdx.x * kWaves + wib; task < ntasks; task += stride) {
    unsigned const v = (task << 6) + lane;
    bool const inr = v < p.n;
    vT old{};
    unsigned r0 = 0, deg = 0;
    if (inr) {
      old = p.V[v];
      vT g = old;
      p.fi(vs(sm.acc[t]), g);
      r0 = p.R[v];
      deg = p.R[v + 1] - r0;
      la.active += 1;
      if constexpr (MIN_FAST) {
        if (is_settled(p.sched, old)) deg = 0;   // nothing can lower it
      }
    }
    la.work += deg;
    unsigned const incl = wave_incl_scan(deg);
    unsigned const total = readlane(incl, 63);
    wave_lds_sync();
    excl[lane] = incl - deg;
    base[lane] = r0;
    wave_lds_sync();
    for (unsigned c0 = 0; c0 < total; c0 += 64) {
      unsigned const e = c0 + lane;
      bool const valid = e < total;
      unsigned own = 64 + lane;
      bool has = false;
      if (valid) {
        own = owner_search(excl, e);
        unsigned const idx = base[own] + (e - excl[own]);
        unsigned const s = p.C[idx];
        if (test_bit(c.in_bm, s)) {
          la.gathers += 1;
          p.fn(vs(sm.tmp[t]), p.V[s], p.E + idx);
          has = true;
        }
      }
      if constexpr (MIN_FAST) {
        if (has) atomic_min_any(&sm.acc[(wib << 6) + own], sm.tmp[t]);
      } else {
        if (ballot64(has)) {
          bool const h2 = seg_reduce(p, sm, own, has);
          unsigned const kprev = __shfl_up(own, 1, 64);
          bool const head = valid && h2 && (lane == 0 || kprev != own);
          if (head) p.fr(vs(sm.acc[(wib << 6) + own]), vs(sm.tmp[t]));
        }
      }
      wave_lds_sync();
    }
    bool upd = false;
    if (inr) {
      upd = p.fu(vs(sm.acc[t]), old);
      if (upd) p.V[v] = sm.acc[t];
    }
    publish_word(p, c, la, upd, v, task);
  }
}

template <bool MIN_FAST, class P, class vT>
__device__ void run_pull_block(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  unsigned const lane = lane_id(), wib = threadIdx.x >> 6, t = threadIdx.x;
  unsigned const ntasks = (p.n + 63) >> 6;
  for (unsigned task = blockIdx.x; task < ntasks; task += c.grid) {
    if (t == 0) sm.ubits = 0;
    for (unsigned j = 0; j < 64; ++j) {
      unsigned const v = (task << 6) + j;
      if (v >= p.n) break;
      unsigned const e0 = p.R[v];
      unsigned e1 = p.R[v + 1];
      if constexpr (MIN_FAST) {
        if (is_settled(p.sched, p.V[v])) e1 = e0;   // nothing can lower it
      }
      bool has = false;
      if (t == 0) la.active += 1;
      for (unsigned e = e0 + t; e < e1; e += blockDim.x) {
        la.work += 1;
        unsigned const s = p.C[e];
        if (test_bit(c.in_bm, s)) {
          la.gathers += 1;
          p.fn(vs(sm.rhs[t]), p.V[s], p.E + e);
          if (has) {
            p.fr(vs(sm.tmp[t]), vs(sm.rhs[t]));
          } else {
            sm.tmp[t] = sm.rhs[t];
            has = true;
          }
        }
      }
      // wave reduction towards lane 0 (partials in tmp[t])
#pragma unroll 1
      for (unsigned s = 32; s > 0; s >>= 1) {
        vT const y = shfl_down_any(sm.tmp[t], s);
        unsigned const hy = __shfl_down((unsigned)has, s, 64);
        if (lane < s && hy) {
          if (has) {
            sm.rhs[t] = y;
            p.fr(vs(sm.tmp[t]), vs(sm.rhs[t]));
          } else {
            sm.tmp[t] = y;
            has = true;
          }
        }
      }
      if (lane == 0) {
        sm.wpart[wib] = sm.tmp[t];
        sm.whas[wib] = has;
      }
      __syncthreads();
      if (t == 0) {
        vT const old = p.V[v];
        vT g = old;
        p.fi(vs(sm.acc[0]), g);
        for (int w = 0; w < kWaves; ++w)
          if (sm.whas[w]) {
            sm.rhs[0] = sm.wpart[w];
            p.fr(vs(sm.acc[0]), vs(sm.rhs[0]));
          }
        if (p.fu(vs(sm.acc[0]), old)) {
          p.V[v] = sm.acc[0];
          sm.ubits |= 1ull << j;
        }
      }
      __syncthreads();
    }
    unsigned long long const ub = sm.ubits;
    if (wib == 0) publish_word(p, c, la, ((ub >> lane) & 1ull) != 0, (task << 6) + lane, task);
    __syncthreads();
  }
}

// ---- edge-balanced (LB) variants -------------------------------------------
//
// The edge space of the iteration's work list (pull / init: every vertex's
// in-edges, offsets R; push: the out-edges of the queue entries, offsets
// written at append time) is cut into equal slices, one per wave of the
// grid, so a hub's edges are spread over many waves.  A wave walks its slice
// in windows of 64 consecutive work items: the 65 offsets of the window go
// to LDS and each lane finds its edge's item with a 6-step search.  Partial
// results of one item within a window are combined in LDS and applied to
// the item's value with one atomic relax per (item, window).

// Largest i in [0, k] with g(i) <= a, where g(i) = off[i] (PLUS = false)
// or off[i] + i (PLUS = true); g non-decreasing, g(0) <= a.  64-ary search:
// every step narrows the range 64x with one load per lane.
template <bool PLUS = false>
__device__ __forceinline__ unsigned lb_find(const unsigned* off, unsigned k, unsigned long long a) {
  unsigned const lane = lane_id();
  unsigned lo = 0, hi = k;  // answer in [lo, hi]
  while (hi - lo > 64) {
    unsigned const span = hi - lo;
    unsigned const idx =
        lo + static_cast<unsigned>((static_cast<unsigned long long>(span) * (lane + 1)) / 65ull);
    bool const pr = static_cast<unsigned long long>(off[idx]) + (PLUS ? idx : 0u) <= a;
    unsigned const cnt = static_cast<unsigned>(__builtin_popcountll(ballot64(pr)));
    unsigned const nlo = cnt ? readlane(idx, cnt - 1) : lo;
    unsigned const nhi = cnt < 64 ? readlane(idx, cnt) - 1 : hi;
    lo = nlo;
    hi = nhi;
  }
  unsigned const idx = lo + 1 + lane;
  bool const pr =
      idx <= hi && static_cast<unsigned long long>(off[idx]) + (PLUS ? idx : 0u) <= a;
  return lo + static_cast<unsigned>(__builtin_popcountll(ballot64(pr)));
}

// Edges that precede position d of the merged sequence of the work list
// (item i's marker at off[i] + i, followed by its edges): d minus the
// number of markers before d.
// Edge positions are relative to base = off[0].
__device__ __forceinline__ unsigned long long merged_to_edge(const unsigned* off, unsigned k,
                                                             unsigned long long d, unsigned base) {
  if (d == 0) return 0;
  unsigned long long const markers =
      static_cast<unsigned long long>(lb_find<true>(off, k, d - 1 + base)) + 1;
  return d - markers;
}

enum lb_kind { LB_PUSH = 0, LB_PULL = 1, LB_INIT = 2 };

template <int KIND, bool MIN_FAST, class P, class vT>
__device__ void run_lb(P& p, iter_ctx& c, lane_acc& la, step_smem<vT>& sm) {
  using eT = typename std::remove_reference<decltype(*p.E)>::type;
  unsigned const lane = lane_id(), wib = threadIdx.x >> 6, t = threadIdx.x;
  // Work list: k items with offsets off[0..k].  Push: the queue (items are
  // queue entries).  Pull and init: the vertices that have in-edges (a
  // static list, so slices never walk runs of edgeless vertices).
  const unsigned* off = (KIND == LB_PUSH) ? c.in_qe : p.nzR;
  unsigned k = (KIND == LB_PUSH) ? c.in_count : p.nnz;
  const unsigned* nzi = p.nzi;
  const unsigned* Cp = p.C;
  eT* Ep = p.E;
  unsigned waves = c.grid * kWaves;
  unsigned wave = blockIdx.x * kWaves + wib;
  if constexpr (KIND == LB_PULL) {
    if (c.split && (c.grid % kSplitParts) == 0) {
      // this workgroup's part of the source-split in-CSR
      unsigned const part = blockIdx.x % kSplitParts;
      unsigned const s0 = p.xseg[part];
      off = p.xnzR + s0;
      nzi = p.xnzi + s0;
      k = p.xseg[part + 1] - s0;
      Cp = p.xC;
      Ep = p.xE;
      waves = (c.grid / kSplitParts) * kWaves;
      wave = (blockIdx.x / kSplitParts) * kWaves + wib;
    }
  }
  if (k == 0) return;
  // Edge positions of the work list run over [base, T) (a push queue's run
  // over [0, its out-edge total)).
  bool const queue = KIND == LB_PUSH;
  unsigned const base = queue ? 0u : off[0];
  unsigned const T = queue ? static_cast<unsigned>(c.in_edges) : off[k];
  if (T == base) return;
  // Merge-path partition: every wave gets an equal share of items + edges,
  // so waves over many low-degree items take fewer edges.
  unsigned long long const W = waves;
  unsigned long long const D = static_cast<unsigned long long>(T - base) + k;
  unsigned long long const per = (D + W - 1) / W;
  unsigned long long const d0 = static_cast<unsigned long long>(wave) * per;
  if (d0 >= D) return;
  unsigned long long const d1 = d0 + per < D ? d0 + per : D;
  unsigned const a = base + static_cast<unsigned>(merged_to_edge(off, k, d0, base));
  unsigned const b = d1 == D ? T : base + static_cast<unsigned>(merged_to_edge(off, k, d1, base));
  if (a >= b) return;
  unsigned* const win = sm.win[wib];
  unsigned* const wbase = sm.base[wib];
  unsigned* const whas = sm.hasf[wib];
  unsigned* const won = sm.excl[wib];
  vT* const wval = sm.val[wib];
  vT* const wslot = sm.slot[wib];
  vT vmax{};
  if constexpr (min_ops<vT>::ok) vmax = std::numeric_limits<vT>::max();

  unsigned cur = lb_find(off, k, a);
  unsigned cpos = a;
  while (cpos < b) {
    // ---- load the window of items [cur, cur + 64) ----
    unsigned const it = cur + lane;
    bool const in_list = it < k;
    wave_lds_sync();
    win[lane] = off[in_list ? it : k];
    if (lane == 0) win[64] = off[(cur + 64 < k) ? cur + 64 : k];
    bool item_on = in_list;
    unsigned vert = 0;
    if constexpr (KIND == LB_PUSH) {
      if (in_list) {
        unsigned const u = c.in_q[it];
        wval[lane] = p.V[u];
        wbase[lane] = p.oR[u];
        la.active += 1;
      }
    } else {
      if (in_list) vert = nzi[it];
      if constexpr (MIN_FAST) {
        // Upper bound of the item's value for the flush's filter (values
        // only decrease), and its out-degree, fetched with the window.
        wval[lane] = in_list ? p.V[vert] : vT{};
      }
      wbase[lane] = in_list ? (p.oR[vert + 1] - p.oR[vert]) : 0u;
      if constexpr (KIND == LB_INIT) {
        item_on = in_list && test_bit(c.in_bm, vert);
        won[lane] = item_on;
      }
      if constexpr (MIN_FAST) wslot[lane] = vmax;
      whas[lane] = 0;
    }
    wave_lds_sync();
    unsigned const wend = win[64];
    unsigned const lim = wend < b ? wend : b;
    bool const skip = (KIND == LB_INIT) && !ballot64(item_on);
    if (!skip) {
      for (unsigned c0 = cpos; c0 < lim; c0 += kChunk) {
        // Lane-consecutive edges: lane l takes [c0 + l*IPL, c0 + l*IPL + IPL).
        // Loads are issued unconditionally (invalid or missing lanes read a
        // clamped, valid address) so they overlap instead of each waiting
        // in its own branch.
        unsigned const eb = c0 + lane * kIPL;
        unsigned const last = lim - 1;
        unsigned e[kIPL], j[kIPL];
        bool valid[kIPL];
        j[0] = owner_search(win, eb < lim ? eb : last);
#pragma unroll
        for (unsigned i = 0; i < kIPL; ++i) {
          e[i] = eb + i;
          valid[i] = e[i] < lim;
          if (i > 0) {
            unsigned jj = j[i - 1];
            unsigned const ec = valid[i] ? e[i] : last;
            while (jj < 63 && win[jj + 1] <= ec) ++jj;
            j[i] = jj;
          }
          la.work += valid[i];
        }
        if constexpr (KIND == LB_PUSH) {
          unsigned dst[kIPL], idx[kIPL];
          bool newly[kIPL];
#pragma unroll
          for (unsigned i = 0; i < kIPL; ++i) {
            unsigned const ec = valid[i] ? e[i] : last;
            idx[i] = wbase[j[i]] + (ec - win[j[i]]);
          }
#pragma unroll
          for (unsigned i = 0; i < kIPL; ++i) dst[i] = p.oC[idx[i]];
          if constexpr (MIN_FAST) {
            // Candidates first, then plain reads of the destinations, then
            // atomics only where the candidate improves on what was read.
            vT cand[kIPL], pre[kIPL];
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) pre[i] = p.V[dst[i]];
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) {
              if (valid[i]) {
                p.fn(vs(sm.tmp[t]), wval[j[i]], p.oE + idx[i]);
                cand[i] = sm.tmp[t];
              }
            }
            // Non-returning atomics: a candidate that beats the plain read
            // counts as a possible move (see relax_push).
            bool imp[kIPL], far[kIPL];
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) {
              imp[i] = valid[i] && cand[i] < pre[i];
              if (imp[i]) atomic_min_any(p.V + dst[i], cand[i]);
            }
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) {
              far[i] = false;
              newly[i] = imp[i] && has_out(p, dst[i]) && claim_next(c, la, dst[i], cand[i], far[i]);
            }
            if (c.nf) {
#pragma unroll
              for (unsigned i = 0; i < kIPL; ++i) far_append(c, far[i], dst[i]);
            }
          } else {
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) {
              newly[i] = false;
              if (valid[i]) {
                p.fn(vs(sm.tmp[t]), wval[j[i]], p.oE + idx[i]);
                if (relax<false>(p, dst[i], sm.tmp[t], sm.rhs[t]) && has_out(p, dst[i]))
                  newly[i] = set_bit_once(c.out_bm, dst[i]);
              }
            }
          }
          stage_queue(p, c, la, newly, dst);
        } else {
          unsigned src[kIPL];
          bool hit[kIPL];
#pragma unroll
          for (unsigned i = 0; i < kIPL; ++i) src[i] = Cp[valid[i] ? e[i] : last];
          if constexpr (KIND == LB_PULL) {
            // a settled destination (its value came with the window)
            // gathers nothing
            bool live[kIPL];
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) {
              live[i] = valid[i];
              if constexpr (MIN_FAST) live[i] = live[i] && !is_settled(p.sched, wval[j[i]]);
            }
            if (c.dense_in && !p.sched.settled_below) {
              // Gathering from a source that did not move is a harmless
              // extra relaxation; it saves the bitmap load of every edge.
              // Not in settled runs: a destination whose in-edges are split
              // over items could be written from a source updated in this
              // iteration (one level too deep) before the item holding its
              // updated in-neighbour sees it settled.
#pragma unroll
              for (unsigned i = 0; i < kIPL; ++i) hit[i] = live[i];
            } else {
              unsigned long long wd[kIPL];
#pragma unroll
              for (unsigned i = 0; i < kIPL; ++i) wd[i] = live[i] ? c.in_bm[src[i] >> 6] : 0ull;
#pragma unroll
              for (unsigned i = 0; i < kIPL; ++i)
                hit[i] = live[i] && ((wd[i] >> (src[i] & 63)) & 1ull);
            }
          } else {
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) hit[i] = valid[i] && won[j[i]] != 0;
          }
          vT sv[kIPL];
          eT ew[kIPL];
#pragma unroll
          for (unsigned i = 0; i < kIPL; ++i) {
            sv[i] = p.V[hit[i] ? src[i] : 0u];
            // Touch the weight so the functor's own read of *connE hits
            // the cache instead of starting a new round trip.
            ew[i] = Ep[hit[i] ? e[i] : 0u];
          }
#pragma unroll
          for (unsigned i = 0; i < kIPL; ++i) keep_live(ew[i]);
          if constexpr (MIN_FAST) {
            // Pre-reduce the lane's consecutive edges of one item in a
            // register, then one LDS atomic per (lane, item).
            unsigned curj = j[0];
            bool has = false;
            vT acc{};
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) {
              if (j[i] != curj) {
                if (has) {
                  atomic_min_any(&wslot[curj], acc);
                  whas[curj] = 1;
                }
                has = false;
                curj = j[i];
              }
              if (hit[i]) {
                la.gathers += 1;
                p.fn(vs(sm.tmp[t]), sv[i], Ep + e[i]);
                vT const cv = sm.tmp[t];
                acc = (has && acc < cv) ? acc : cv;
                has = true;
              }
            }
            if (has) {
              atomic_min_any(&wslot[curj], acc);
              whas[curj] = 1;
            }
          } else {
#pragma unroll
            for (unsigned i = 0; i < kIPL; ++i) {
              if (hit[i]) {
                la.gathers += 1;
                p.fn(vs(sm.tmp[t]), sv[i], Ep + e[i]);
              }
              if (ballot64(hit[i])) {
                unsigned const key = valid[i] ? j[i] : 64 + lane;
                bool const h2 = seg_reduce(p, sm, key, hit[i]);
                unsigned const kprev = __shfl_up(key, 1, 64);
                bool const head = valid[i] && h2 && (lane == 0 || kprev != key);
                if (head) {
                  if (whas[j[i]]) {
                    p.fr(vs(wslot[j[i]]), vs(sm.tmp[t]));
                  } else {
                    wslot[j[i]] = sm.tmp[t];
                    whas[j[i]] = 1;
                  }
                }
                wave_lds_sync();
              }
            }
          }
          wave_lds_sync();
        }
      }
      if constexpr (KIND != LB_PUSH) {
        // Apply each item's window partial to its value.
        wave_lds_sync();
        bool newly = false;
        if (in_list && whas[lane]) {
          bool moved;
          if constexpr (MIN_FAST) {
            vT const pv = wslot[lane];
            moved = pv < wval[lane] && pv < atomic_min_any(p.V + vert, pv);
          } else {
            sm.tmp[t] = wslot[lane];
            moved = relax<false>(p, vert, sm.tmp[t], sm.rhs[t]);
          }
          if (moved && wbase[lane] != 0) {
            unsigned long long const bit = 1ull << (vert & 63);
            newly = !(atomicOr(c.out_bm + (vert >> 6), bit) & bit);
          }
        }
        append_queue(p, c, la, newly, vert);
      }
    }
    cpos = lim;
    cur += 64;
  }
}

// ---- scheduling of the next iteration (last workgroup, one wave) ----------
__device__ __forceinline__ unsigned plain_variant(const sched_params& s, unsigned long long sum_outdeg,
                                                  unsigned long long m) {
  bool const pull = static_cast<double>(sum_outdeg) * static_cast<double>(s.alpha) >
                    static_cast<double>(m);
  // a settled run's pull is the one that stops at the first updated
  // in-neighbour (thread per vertex)
  if (pull && s.settled_below) return V_PULL_T;
  return (pull ? V_PULL_T : V_PUSH_T) + s.fixed_gran;
}

// Variant of iteration `next_iter`, from the statistics of the updated set
// it will consume.
// Called by the 64 lanes of one wave.
__host__ __device__ inline unsigned quarter_octave(unsigned long long x) {
  // floor(4 * log2(1 + x)) from the leading bit and the next two bits
  unsigned long long const y = x + 1;
  unsigned const lb = 63u - static_cast<unsigned>(__builtin_clzll(y));
  unsigned const frac = lb >= 2 ? static_cast<unsigned>((y >> (lb - 2)) & 3u)
                                : static_cast<unsigned>((y << (2 - lb)) & 3u);
  return 4u * lb + frac;
}

__host__ __device__ inline unsigned learned_key(unsigned count, unsigned long long sdeg,
                                                unsigned maxdeg, unsigned iter) {
  unsigned const a = quarter_octave(count) & 0xffu;
  unsigned const b = quarter_octave(sdeg) & 0xffu;
  unsigned const c = (quarter_octave(maxdeg) >> 2) & 0xffu;
  unsigned const d = quarter_octave(iter) & 0xffu;
  return (a << 24) | (b << 16) | (c << 8) | d;
}

// A push iteration over a small queue (entries + out-edges <= kSmallWork)
// needs few workgroups: it runs on the first small_grid workgroups of its
// launch, and the host replays the following chunk of launches on a small
// grid.  Not PUSH_B: it runs one queue entry per workgroup and needs many.
__device__ __forceinline__ bool small_iteration(unsigned variant, unsigned count,
                                                unsigned long long sdeg) {
  return (variant == V_PUSH_T || variant == V_PUSH_W || variant == V_PUSH_LB) &&
         static_cast<unsigned long long>(count) + sdeg <= kSmallWork;
}

// The grid level of the chunk after an iteration that chose `variant` for a
// queue of `count` entries with `sdeg` out-edges: 0 full, 1 small, 2 tiny.
__device__ __forceinline__ unsigned grid_level(unsigned variant, unsigned count,
                                               unsigned long long sdeg) {
  if (!small_iteration(variant, count, sdeg)) return 0u;
  return static_cast<unsigned long long>(count) + sdeg <= kTinyWork ? 2u : 1u;
}

// What the learned schedule's decision reads besides the statistics,
// prepared by the caller so that it can load them early: the lanes' network
// weights (mlp_load_lane; with kLds, LDS to stage them in instead) and the
// decision cache's last key and variant from the control line.
struct learned_in {
  mlp_lane_weights w;
  float* scratch = nullptr;
  unsigned cache_key = 0xffffffffu;
  unsigned cache_variant = 0;
};

template <bool kLds = false, class P>
__device__ unsigned choose_variant(P& p, unsigned count, unsigned long long sdeg, unsigned maxdeg,
                                   unsigned next_iter, learned_in const& li) {
  switch (p.sched.mode) {
    case SCHED_PLAIN:
      return plain_variant(p.sched, sdeg, p.m);
    case SCHED_FORCED:
      return p.sched.forced_variant;
    case SCHED_LEARNED: {
      unsigned const fb = plain_variant(p.sched, sdeg, p.m);
      // Settled runs (BFS) take the plain rule, a direction-optimizing
      // switch: on the recorded settled states it is within 0.6 % of the
      // fastest variant's time (the network's choices: 0.9 %).
      if (p.sched.settled_below) return fb;
      // Small-work guard: a queue whose entries + out-edges fit kSmallWork
      // costs about the fixed cost of an iteration under any push variant
      // on the small grid, so the plain choice stands and the inference is
      // skipped (long runs of small frontiers: high-diameter graphs).
      if (static_cast<unsigned long long>(count) + sdeg <= kSmallWork) return fb;
      // Consecutive iterations with the same quantized features (a quarter
      // octave of frontier size and out-edges, an octave of max degree and
      // a quarter octave of the iteration index) reuse the last decision.
      unsigned const key = learned_key(count, sdeg, maxdeg, next_iter);
      unsigned v;
      if (key == li.cache_key) {
        v = li.cache_variant;
      } else {
        float feat[MLP_IN];
        make_features(feat, p.n, p.m, count, sdeg, maxdeg, next_iter, p.sched.settled_below != 0);
        if constexpr (kLds)
          v = mlp_choose_lds(p.sched.mlp, feat, fb, li.scratch, p.sched.margin);
        else
          v = mlp_choose_lane(p.sched.mlp, li.w, feat, fb, p.sched.margin);
        if (lane_id() == 0) {
          p.ctl->cache_key = key;
          p.ctl->cache_variant = v;
        }
      }
      // Cost guard: a pull iteration reads every in-edge, so it cannot win
      // when the frontier's out-edges are a small fraction of m.
      unsigned long long const guard = p.sched.settled_below ? kPullGuardSettled : kPullGuard;
      if (v >= V_PULL_T && v <= V_PULL_LB && sdeg * guard < p.m) return fb;
      // A settled run's pull is the thread-per-vertex one (the only pull
      // that stops at the first updated in-neighbour), as in the plain rule.
      if (v >= V_PULL_T && v <= V_PULL_LB && p.sched.settled_below) return V_PULL_T;
      return v;
    }
    default:
      return V_CLASSIC;
  }
}

template <class P, class vT>
__device__ void finalize_iteration(P& p, iter_ctx& c, unsigned iter, step_smem<vT>& sm) {
  // Every statistic and queue counter of this workgroup was updated by an
  // agent-scope atomic whose result has returned (or, for thread 0's own
  // counter updates, is waited for here) before the ticket is taken, so the
  // last workgroup reads complete counters without an L2 write-back fence.
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_s_waitcnt(0);
    sm.is_last = take_ticket(p.board);
  }
  __syncthreads();
  if (!sm.is_last) return;
  if (threadIdx.x >= 64) return;
  step_stats* st = p.stats + iter;
  bool const learned = p.sched.mode == SCHED_LEARNED;
  unsigned const ng = gridDim.x < kTicketGroups ? gridDim.x : kTicketGroups;
  // The queue counter, the iteration limit, the learned schedule's
  // max-degree partials, decision cache and network weights are all loaded
  // before any of them is used, so they come back in one round trip (the
  // fold and the inference would otherwise each start another one on this
  // serial path).  Both queue counters are read whatever the variant;
  // lanes past the group count read slot 0 again (max is unchanged).
  unsigned long long const qe =
      __hip_atomic_load(&st->qe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned const scount =
      __hip_atomic_load(&st->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned const max_iter = p.ctl->max_iter;
  bool const pre_md = learned && !c.no_deg && !p.collect && !c.nf;
  unsigned const md_part =
      __hip_atomic_load(&p.gst[lane_id() < ng ? lane_id() : 0u].max_deg, __ATOMIC_RELAXED,
                        __HIP_MEMORY_SCOPE_AGENT);
  // (no branch between the loads: a run without a network reads the group
  // slots as its weights, any readable MLP_NPARAM floats, and never uses
  // them)
  static_assert(sizeof(group_stats) * kTicketGroups >= sizeof(float) * MLP_NPARAM,
                "the group slots stand in for absent weights");
  learned_in li;
  li.cache_key = p.ctl->cache_key;
  li.cache_variant = p.ctl->cache_variant;
  mlp_load_lane(p.sched.mlp ? p.sched.mlp : reinterpret_cast<const float*>(p.gst), li.w);
  unsigned count;
  unsigned long long sdeg;
  if (c.scatter) {
    count = scount;
    sdeg = 0;
  } else {
    // Queue length and out-edge total from the packed append counter; the
    // sentinel offset closes the queue's offset array.
    count = qe_count(qe);
    sdeg = qe_edges(qe);
    if (threadIdx.x == 0) {
      st->count = count;
      st->sum_outdeg = sdeg;
      c.out_qe[count] = static_cast<unsigned>(sdeg);
    }
  }
  unsigned maxdeg = 0;
  unsigned long long vmin = ~0ull;
  // The learned schedule uses the next frontier's max degree only when its
  // small-work guard does not decide (choose_variant); the group slots are
  // cleared either way (an iteration that skipped the statistics left them
  // at zero).
  bool const guarded = static_cast<unsigned long long>(count) + sdeg <= kSmallWork ||
                       p.sched.settled_below != 0;
  if (pre_md) {
    if (!guarded) maxdeg = wave_max(md_part);
    if (lane_id() < ng) p.gst[lane_id()].max_deg = 0;
  } else if (p.collect || c.nf || (learned && !guarded)) {
    maxdeg = fold_group_stats(p.gst, st, p.collect != 0, c.nf, vmin);
  }
  // Near-far: the far pile's size and least value after this iteration; a
  // queue that ran dry while the pile holds entries makes the next
  // iteration a split at threshold (least far value + delta).
  bool split_next = false;
  unsigned far_n = 0;
  unsigned long long far_vmin = ~0ull;
  if (c.nf) {
    if (c.variant == V_SPLIT) {
      far_n = __hip_atomic_load(&p.ctl->far_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      far_vmin = vmin;   // exact: the least value of the entries kept
    } else {
      far_n = __hip_atomic_load(c.far_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      unsigned long long const fv = p.ctl->far_vmin;
      far_vmin = vmin < fv ? vmin : fv;
    }
    split_next = count == 0 && far_n != 0;
  }
  unsigned next = V_CLASSIC;
  bool const stop = (count == 0 && !split_next) || (iter + 1 >= max_iter);
  if (!stop)
    next = split_next ? static_cast<unsigned>(V_SPLIT)
                      : choose_variant(p, count, sdeg, maxdeg, iter + 1, li);
  if (threadIdx.x == 0) {
    p.ctl->next_variant = next;
    p.ctl->iter = iter + 1;
    if (c.nf) {
      if (c.variant == V_SPLIT) {
        p.ctl->far_sel ^= 1u;
        p.ctl->far_n = far_n;
        p.ctl->far_next = 0;
      }
      p.ctl->far_vmin = far_vmin;
      if (split_next) {
        unsigned long long const t = far_vmin + p.sched.delta;
        p.ctl->thresh_prev = c.T;
        p.ctl->thresh = t < far_vmin ? ~0ull : t;
      }
    }
    p.ctl->small_next = stop         ? 0u
                        : split_next ? (far_n <= kSmallWork ? 1u : 0u)
                                     : grid_level(next, count, sdeg);
    p.ctl->q_count = count;
    p.ctl->q_edges = sdeg;
    if (stop) {
      p.ctl->done = 1;
      if (count != 0 || split_next) p.ctl->overflow = 1;
    }
  }
}

// DELTA = false compiles only CLASSIC (any functor set); DELTA = true also
// compiles the push/pull variants (min-semilattice functor sets on 4- or
// 8-byte values).  MIN_FAST additionally replaces red/upd by native min and
// compare (functor sets whose red is min and upd is "<" on an integer type).
template <bool DELTA, bool MIN_FAST, bool NF, class P>
__global__ __launch_bounds__(COMPILE_TIME_DETERMINED_BLOCK_SIZE) void step_kernel(P p) {
  using vT = typename std::remove_reference<decltype(*p.V)>::type;
  __shared__ step_smem<vT> sm;
  unsigned const iter = p.ctl->iter;
  if (p.ctl->done) return;
  unsigned const variant = p.ctl->next_variant;
  iter_ctx c;
  c.iter = iter;
  unsigned long long const nw = p.nwords;
  c.in_bm = p.bm + (iter % 3) * nw;
  c.out_bm = p.bm + ((iter + 1) % 3) * nw;
  c.clr_bm = p.bm + ((iter + 2) % 3) * nw;
  c.in_q = p.q + (iter & 1) * static_cast<unsigned long long>(p.n);
  c.out_q = p.q + ((iter + 1) & 1) * static_cast<unsigned long long>(p.n);
  c.in_qe = p.qe + (iter & 1) * (static_cast<unsigned long long>(p.n) + 1);
  c.out_qe = p.qe + ((iter + 1) & 1) * (static_cast<unsigned long long>(p.n) + 1);
  // The input queue's length and out-edge total: from the control line
  // (the previous iteration wrote them there), or for iteration 0 the
  // seeded queue and its sentinel offset (push variants only).
  unsigned long long in_edges = 0;
  if (iter == 0) {
    c.in_count = p.ctl->seed_count;
    if (variant >= V_PUSH_T && variant <= V_PUSH_LB) in_edges = c.in_qe[c.in_count];
  } else {
    c.in_count = p.ctl->q_count;
    in_edges = p.ctl->q_edges;
  }
  c.in_edges = in_edges;
  c.tiny = static_cast<unsigned long long>(c.in_count) + in_edges <= kTinyWork;
  c.dense_in = iter > 0 && in_edges * 10ull >= p.m * 7ull;
  // The split pays where the pull gathers much (the updated set's out-edges
  // cover a quarter of the edges or more); on a sparse pull its extra
  // segments cost more than the local gathers save.
  c.split = p.xsplit && iter > 0 && in_edges * 4ull >= p.m;
  // The learned schedule's max-degree statistics serve only decisions its
  // small-work guard does not take.  A push iteration appends at most one
  // entry per out-edge it expands, each with at most max_outdeg out-edges,
  // so when out-edges(queue) * (1 + max_outdeg) <= kSmallWork the next
  // decision is guarded and the statistics (a returning atomic per
  // workgroup and the slot clear) are skipped.
  // Settled runs never infer (they take the plain rule): no statistics.
  c.no_deg = p.sched.mode == SCHED_LEARNED && !p.collect && !(NF && MIN_FAST) &&
             (p.sched.settled_below != 0 ||
              (iter > 0 && p.max_outdeg != 0 && variant >= V_PUSH_T && variant <= V_PUSH_LB &&
               in_edges <= kSmallWork / (1ull + p.max_outdeg)));
  c.variant = variant;
  c.nf = false;
  c.T = ~0ull;
  c.Told = 0;
  c.far_q = nullptr;
  c.far_src = nullptr;
  c.far_bm = nullptr;
  c.far_n = nullptr;
  if constexpr (NF && MIN_FAST && std::is_unsigned<vT>::value) {
    c.nf = true;
    c.T = p.ctl->thresh;
    c.Told = p.ctl->thresh_prev;
    unsigned long long const sel = p.ctl->far_sel & 1u;
    c.far_q = p.fq + sel * static_cast<unsigned long long>(p.n);
    c.far_bm = p.fbm;
    c.far_n = &p.ctl->far_n;
    if (variant == V_SPLIT) {
      // the split reads the current pile and appends the entries it keeps
      // to the other one
      c.in_count = p.ctl->far_n;
      c.far_src = c.far_q;
      c.far_q = p.fq + (sel ^ 1ull) * static_cast<unsigned long long>(p.n);
      c.far_n = &p.ctl->far_next;
    }
  }
  c.st = p.stats + iter;
  c.scatter = (p.sched.mode == SCHED_CLASSIC);
  // A push iteration runs on the first workgroups only, as many as its work
  // needs (kWorkPerGroup; the queue's out-edge total is the sentinel offset
  // c.in_edges).  The others skip the work but still take the
  // completion ticket: the last workgroup must be the last of the whole grid
  // to start, or a late one could read the control block of the next
  // iteration.
  c.grid = gridDim.x;
  if (p.small_grid && p.small_grid < gridDim.x &&
      (variant == V_PUSH_T || variant == V_PUSH_W || variant == V_PUSH_LB ||
       variant == V_SPLIT)) {
    unsigned long long const work =
        static_cast<unsigned long long>(c.in_count) +
        (variant == V_SPLIT ? 0ull : in_edges);
    unsigned long long const want = (work + kWorkPerGroup - 1) / kWorkPerGroup;
    c.grid = want <= p.small_grid ? p.small_grid
             : want < gridDim.x   ? static_cast<unsigned>(want)
                                  : gridDim.x;
  }
  bool const idle = blockIdx.x >= c.grid;
  unsigned const wib = threadIdx.x >> 6;
  c.abuf_v = sm.abuf_v[wib];
  c.abuf_e = sm.abuf_e[wib];
  c.ahdr = sm.ahdr[wib];
  c.acap = c.nf ? kAcap / 2 : kAcap;
  c.fbuf = sm.abuf_v[wib] + kAcap / 2;
  if ((threadIdx.x & 63) == 0) {
    sm.ahdr[wib][0] = 0;
    sm.ahdr[wib][1] = 0;
    sm.ahdr[wib][2] = 0;
  }
  if (threadIdx.x == 0) {
    sm.bstat[0] = sm.bstat[1] = sm.bstat[2] = 0;
    sm.bvmin = ~0ull;
    sm.bmax = 0;
  }
  __syncthreads();
  lane_acc la;
  if (idle) {
    // counts in only
  } else if constexpr (DELTA) {
    switch (variant) {
      case V_CLASSIC: run_classic<MIN_FAST>(p, c, la, sm); break;
      case V_PUSH_T: run_push_thread<MIN_FAST>(p, c, la, sm); break;
      case V_PUSH_W: run_push_wave<MIN_FAST>(p, c, la, sm); break;
      case V_PUSH_B: run_push_block<MIN_FAST>(p, c, la, sm); break;
      case V_PULL_T: run_pull_thread<MIN_FAST>(p, c, la, sm); break;
      case V_PUSH_LB: run_lb<LB_PUSH, MIN_FAST>(p, c, la, sm); break;
      case V_PULL_W: run_pull_wave<MIN_FAST>(p, c, la, sm); break;
      case V_PULL_B: run_pull_block<MIN_FAST>(p, c, la, sm); break;
      case V_PULL_LB: run_lb<LB_PULL, MIN_FAST>(p, c, la, sm); break;
      case V_SPLIT:
        if constexpr (NF) run_split(p, c, la, sm);
        break;
      default: run_lb<LB_INIT, MIN_FAST>(p, c, la, sm); break;
    }
  } else {
    run_classic<false>(p, c, la, sm);
  }
  if (!idle) {
    bool const want_stats =
        p.collect || (p.sched.mode == SCHED_LEARNED && !c.no_deg) || c.nf;
    if (want_stats) flush_stats(sm, la, p.collect != 0, c.nf);
    block_flush(c, sm, want_stats ? p.gst : nullptr);
    // Clear the bitmap that iteration iter+2 will write.
    for (unsigned w = blockIdx.x * blockDim.x + threadIdx.x; w < p.nwords; w += c.grid * blockDim.x)
      c.clr_bm[w] = 0ull;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.st->variant = variant;
    c.st->in_count = c.in_count;
  }
  finalize_iteration(p, c, iter, sm);
}

// Zeroes `count` device words (a kernel, so that no memset node is needed).
static __global__ void zero_words_kernel(unsigned* __restrict__ p, unsigned count) {
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Resets the control block, the statistics and all three bitmap slots of a
// run.
static __global__ void reset_kernel(step_stats* stats, unsigned nstats, unsigned long long* bm_ring,
                             unsigned nwords, step_ctl* ctl,
                             unsigned first_variant, unsigned max_iter, unsigned* update_flag,
                             group_stats* gst, unsigned* board, unsigned long long delta,
                             unsigned long long* far_bm) {
  unsigned const tid = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned const nt = gridDim.x * blockDim.x;
  if (tid < kTicketGroups) {
    gst[tid] = group_stats{};
    gst[tid].vmin = ~0ull;
  }
  clear_ticket_board(board, tid, nt);
  unsigned long long* sw = reinterpret_cast<unsigned long long*>(stats);
  unsigned const nsw = nstats * (sizeof(step_stats) / 8);
  for (unsigned i = tid; i < nsw; i += nt) sw[i] = 0ull;
  for (unsigned i = tid; i < 3 * nwords; i += nt) bm_ring[i] = 0ull;
  if (far_bm)
    for (unsigned i = tid; i < nwords; i += nt) far_bm[i] = 0ull;
  if (tid == 0) {
    ctl->iter = 0;
    ctl->done = 0;
    ctl->next_variant = first_variant;
    ctl->max_iter = max_iter;
    ctl->overflow = 0;
    ctl->seed_count = 0;
    ctl->small_next = 0;
    ctl->q_count = 0;
    ctl->q_edges = 0;
    ctl->cache_key = 0xffffffffu;
    ctl->cache_variant = 0;
    ctl->thresh = delta ? delta : ~0ull;
    ctl->thresh_prev = 0;
    ctl->far_vmin = ~0ull;
    ctl->far_n = 0;
    ctl->far_next = 0;
    ctl->far_sel = 0;
    if (update_flag) *update_flag = 0;
  }
}

// Seeded start of a delta run: the first iteration propagates from a given
// updated set instead of recomputing an initial frontier.  SEED_SOURCE:
// {source}; SEED_ALL: every vertex (queue = identity, offsets = oR);
// SEED_LIST: `count` distinct vertices from `list` (one workgroup scans
// their out-degrees into the queue offsets).  Runs after reset_kernel;
// picks the first variant with the run's scheduler.
enum seed_mode : unsigned { SEED_NONE = 0, SEED_SOURCE = 1, SEED_ALL = 2, SEED_LIST = 3 };

template <class P>
__global__ void seed_kernel(P p, unsigned mode, unsigned src, const unsigned* list,
                            unsigned list_count) {
  __shared__ unsigned wtot[COMPILE_TIME_DETERMINED_BLOCK_SIZE / WAVE_SIZE];
  __shared__ unsigned long long s_sum;
  __shared__ unsigned s_max;
  unsigned const tid = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned const nt = gridDim.x * blockDim.x;
  if (mode == SEED_ALL) {
    for (unsigned i = tid; i < p.n; i += nt) p.q[i] = i;
    for (unsigned i = tid; i <= p.n; i += nt) p.qe[i] = p.oR[i];
    for (unsigned w = tid; w < p.nwords; w += nt) {
      unsigned const lo = w * 64u;
      unsigned long long word = 0;
      if (lo < p.n) word = (p.n - lo >= 64) ? ~0ull : ((1ull << (p.n - lo)) - 1ull);
      p.bm[w] = word;
    }
  }
  if (mode == SEED_LIST && blockIdx.x == 0) {
    // Block-wide exclusive scan of the out-degrees, 256 entries at a time.
    unsigned const t = threadIdx.x, lane = lane_id(), wib = t >> 6;
    if (t == 0) {
      s_sum = 0;
      s_max = 0;
    }
    __syncthreads();
    for (unsigned base = 0; base < list_count; base += blockDim.x) {
      unsigned const i = base + t;
      unsigned v = 0, d = 0;
      if (i < list_count) {
        v = list[i];
        d = p.oR[v + 1] - p.oR[v];
      }
      unsigned const incl = wave_incl_scan(d);
      if (lane == 63) wtot[wib] = incl;
      __syncthreads();
      unsigned before = 0, total = 0;
      for (unsigned w = 0; w < blockDim.x / WAVE_SIZE; ++w) {
        if (w < wib) before += wtot[w];
        total += wtot[w];
      }
      unsigned const carry = static_cast<unsigned>(s_sum);
      if (i < list_count) {
        p.q[i] = v;
        p.qe[i] = carry + before + incl - d;
        atomicOr(p.bm + (v >> 6), 1ull << (v & 63));
        atomicMax(&s_max, d);
      }
      __syncthreads();
      if (t == 0) s_sum += total;
      __syncthreads();
    }
    if (t == 0) p.qe[list_count] = static_cast<unsigned>(s_sum);
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    unsigned count;
    unsigned long long sdeg;
    unsigned maxdeg = 0;
    if (mode == SEED_ALL) {
      count = p.n;
      sdeg = p.m;
    } else if (mode == SEED_LIST) {
      count = list_count;
      sdeg = s_sum;
      maxdeg = s_max;
    } else {
      unsigned const d = p.oR[src + 1] - p.oR[src];
      if (threadIdx.x == 0) {
        p.q[0] = src;
        p.qe[0] = 0;
        p.qe[1] = d;
        p.bm[src >> 6] |= 1ull << (src & 63);
      }
      count = 1;
      sdeg = d;
      maxdeg = d;
    }
    learned_in li;
    if (p.sched.mode == SCHED_LEARNED && p.sched.mlp) {
      li.cache_key = p.ctl->cache_key;
      li.cache_variant = p.ctl->cache_variant;
      mlp_load_lane(p.sched.mlp, li.w);
    }
    unsigned const v = choose_variant(p, count, sdeg, maxdeg, 0, li);
    if (threadIdx.x == 0) {
      p.ctl->seed_count = count;
      p.ctl->next_variant = v;
    }
  }
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_STEP_KERNEL_HPP_
