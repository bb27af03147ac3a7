// engine.hpp -- process(): the fixpoint driver.
//
// Drop-in for the single-GPU overloads of the reference's process()
// (kites/device/compute.cuh:29-137 and :397-411): same template signature,
// same functor contract, the result left in devGrph.V (or copied back into
// the host graph's V by the host-graph overload), the initial frontier
// consumed.  What differs is how the loop runs on MI355X:
//
//  * No host round trip per iteration.  Each iteration is one launch of
//    step_kernel; the next iteration's variant, the done flag and the
//    iteration index are decided on the device by the last workgroup of the
//    previous launch.  The host replays a captured hipGraph of `chunk`
//    launches and reads the done flag once per chunk (launches after
//    convergence return immediately).
//  * The functor set is probed on the host (the lambdas are __host__
//    __device__).  A set that behaves as a min-semilattice (init copies; red
//    is idempotent, commutative and associative; upd reports whether red
//    moved the value) may run the delta variants after the first iteration
//    (push over CSR from a sparse queue, pull over CSC with a dense bitmap);
//    any other set runs the reference algorithm (CLASSIC) every iteration.
//    For a monotone min-semilattice set the fixpoint does not depend on the
//    variant sequence, so every schedule produces the same vertex values.
//  * Schedules: classic, plain (fixed push/pull threshold and granularity),
//    forced (one variant for every delta iteration; used by the timing
//    recorder) and learned (the MLP picks the variant per iteration, with
//    the plain choice as the fallback).
#ifndef KITES_ENGINE_HPP_
#define KITES_ENGINE_HPP_

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <iostream>
#include <limits>
#include <random>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "detail/ordered_kernel.hpp"
#include "detail/persist_kernel.hpp"
#include "detail/step_kernel.hpp"
#include "device.hpp"
#include "graph_device.hpp"
#include "split.hpp"

namespace kites {

enum class schedule { automatic, classic, plain, forced, learned };
// Granularity of the delta variants: one lane, one wave (merged edge list of
// 64 work items) or one workgroup per work item, or edge-balanced slices.
enum class granularity : unsigned { thread = 0, wave = 1, workgroup = 2, edge_balanced = 3 };

// One state of a run and the time every delta variant takes from it (the
// scheduler's training data): feat = the scheduler features of the
// iteration, ms[k] = device time of variant k + 1 started from that state.
struct variant_sample {
  float feat[detail::MLP_IN];
  float ms[detail::MLP_OUT];
  unsigned iter;
  unsigned plain_variant;
};

// Per-run report (optional output of process()).
struct run_report {
  unsigned iterations = 0;
  double ms = 0.0;             // device time of the iteration loop (hipEvents)
  bool delta = false;          // delta variants were allowed
  bool min_fast = false;       // native min path was used
  unsigned graph_builds = 0;   // iteration-chunk graphs captured by this graph's runs so far
  std::vector<detail::step_stats> stats;  // one entry per iteration (if requested)
};

struct options {
  schedule sched = schedule::automatic;
  granularity gran = granularity::edge_balanced;  // plain schedule's granularity
  unsigned forced_variant = detail::V_PULL_LB;
  float alpha = 14.f;                    // plain: pull when out-edges(frontier)*alpha > m
  float margin = 0.f;                    // learned: min predicted log2 gain over plain's choice
  // Near-far push (unsigned min functor sets): queue entries whose value is
  // at or above a moving threshold are carried instead of expanded; the
  // threshold moves to the least carried value + delta when nothing below
  // it is left.  0 = off.  Results are the same fixpoint.
  unsigned long long delta = 0;
  // Settled values (unsigned min functor sets, delta schedules; 0 = off):
  // the caller guarantees that a value below this bound never changes once
  // written and that every candidate a vertex is offered in the iteration
  // that first lowers it below the bound is the same -- true of BFS levels
  // from sources at 0 with the unreached value as the bound.  Pulls then
  // skip settled vertices, and the thread-per-vertex pull (the plain
  // schedule's pull) stops at the first updated in-neighbour.
  unsigned long long settled_below = 0;
  const float* mlp_weights = nullptr;    // device pointer, learned schedule
  unsigned max_iterations = 1u << 16;
  bool use_graph = true;
  unsigned chunk = 16;
  unsigned grid = 0;                     // 0 = occupancy-derived
  bool adaptive_grid = true;             // small grid for chunks of small push frontiers
  unsigned small_grid = 0;               // workgroups of that grid (0 = one per compute unit)
  unsigned tiny_grid = 0;                // after a tiny push frontier (0 = a quarter of those)
  // After the first chunk, queue the next chunk before reading the last
  // one's control line (needs the workspace's events; not without).
  bool pipeline = true;
  // Small push frontiers (min functor sets, plain and learned schedules,
  // near-far too): the chunks that follow a tiny one start with a
  // one-workgroup loop that runs every iteration whose queue fits it
  // (detail/persist_kernel.hpp) in one launch.  Needs adaptive_grid.  persist_edges: the largest out-edge
  // total of a queue the loop takes (at most detail::kPersistEdges).
  bool persist = true;
  unsigned persist_edges = detail::kPersistDefaultEdges;
  // Edge-balanced pull over the source-split in-CSR (split.hpp): each XCD
  // gathers source values from its own eighth of V.  Holds one more copy of
  // C and E, built on the first run that sets it (delta schedules only).
  bool split_pull = false;
  bool verbose = false;
  bool keep_stats = false;
  run_report* report = nullptr;
  // Ordered mode (any functor set, vertex-per-group 1): Jacobi iterations
  // (each reads the previous iteration's values) with a fixed combine order
  // of every vertex's in-edges (detail/ordered_kernel.hpp), so reductions
  // that are not idempotent, not associative or not commutative -- sums,
  // PageRank-style updates, floating point -- give one defined result.
  bool ordered = false;
  // Seeded start (delta schedules only): instead of recomputing the initial
  // frontier, the first iteration propagates from {seed_source} (seed_source
  // mode) or from every vertex (seed_all mode).  Equivalent to the frontier
  // start whenever the initial values differ from a fixpoint only through
  // those vertices (single-source SSSP/BFS from a source whose frontier is
  // its out-neighbours; min-label CC with every vertex active).
  // seed_list mode: seed_count distinct vertices at seed_list (a device
  // pointer), e.g. one source per graph of a batch of disjoint graphs.
  enum class seeding { none, source, all, list } seed = seeding::none;
  unsigned seed_source = 0;
  const unsigned* seed_list = nullptr;
  unsigned seed_count = 0;
  // Recorder (delta schedules): before every iteration of the run, fork the
  // state and time each delta variant from it; append one sample per
  // iteration.  The run itself continues with the schedule's own choices.
  std::vector<variant_sample>* record = nullptr;
  unsigned record_repeats = 2;
  // 0: time only the forked iteration; else time it plus the rest of the
  // run with this variant forced (time-to-finish of the choice).
  unsigned record_rollout = 0;
};

namespace detail {

struct probe_result {
  bool delta = false;
  bool min_fast = false;
};

// Host-side check of the functor set on sample values (see the header
// comment).  Only 4- and 8-byte trivially copyable value types can run the
// delta variants (they need an atomic compare-and-swap).
template <class vT, class eT, class FI, class FN, class FR, class FU>
probe_result probe_functors(FI fi, FN, FR fr, FU fu) {
  probe_result r;
  if constexpr (!(std::is_trivially_copyable<vT>::value && (sizeof(vT) == 4 || sizeof(vT) == 8) &&
                  std::is_integral<vT>::value)) {
    return r;
  } else {
    // 15 edge values and 48 seeded random ones (a fixed array: the probe
    // allocates nothing)
    std::array<vT, 63> s{};
    vT const mx = std::numeric_limits<vT>::max();
    vT const mn = std::numeric_limits<vT>::min();
    std::size_t k = 0;
    for (vT x : {vT(0), vT(1), vT(2), vT(3), vT(7), vT(64), vT(100), vT(255), vT(12345),
                 vT(mx / 2), vT(mx - 2), vT(mx - 1), mx, mn, vT(mn + 1)})
      s[k++] = x;
    std::mt19937_64 rng(0x6b697465);
    while (k < s.size()) s[k++] = static_cast<vT>(rng());
    bool init_id = true, is_min = true, upd_lt = true, semi = true, upd_moved = true;
    auto red = [&](vT a, vT b) {
      volatile vT l = a, rr = b;
      fr(l, rr);
      return static_cast<vT>(l);
    };
    for (vT a : s) {
      volatile vT loc{};
      vT g = a;
      fi(loc, g);
      if (static_cast<vT>(loc) != a) init_id = false;
      if (red(a, a) != a) semi = false;
      for (vT b : s) {
        vT const ab = red(a, b);
        if (ab != std::min(a, b)) is_min = false;
        if (ab != red(b, a)) semi = false;
        volatile vT nv = a;
        bool const u = fu(nv, b);
        if (u != (a < b)) upd_lt = false;
        vT const rb = red(b, a);
        volatile vT nv2 = rb;
        if (fu(nv2, b) != (rb != b)) upd_moved = false;
      }
    }
    for (std::size_t i = 0; i + 2 < s.size(); i += 3)
      if (red(red(s[i], s[i + 1]), s[i + 2]) != red(s[i], red(s[i + 1], s[i + 2]))) semi = false;
    r.delta = init_id && semi && upd_moved;
    r.min_fast = r.delta && is_min && upd_lt;
    return r;
  }
}

template <bool DELTA, bool MIN_FAST, bool NF, class P>
unsigned step_grid(gpu& dev, unsigned requested) {
  if (requested) return requested;
  int per_cu = 0;
  KITES_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(&step_kernel<DELTA, MIN_FAST, NF, P>),
      COMPILE_TIME_DETERMINED_BLOCK_SIZE, 0));
  per_cu = std::max(1, std::min(per_cu, 8));
  // Three rounds of resident workgroups: later workgroups start on the
  // compute units that finish first, which evens out the edge-balanced
  // slices' uneven latencies.
  return static_cast<unsigned>(dev.numComputeUnits() * per_cu * 3);
}

// Device state of one run -- the 3-slot frontier bitmap ring, the two
// queues and their out-edge offsets, the per-iteration statistics, the
// control block and the per-group statistics -- and the host side of its
// loop.  process() takes it from the graph's workspace; process_noalloc()
// from memory the caller owns.
struct run_state {
  unsigned long long* bm = nullptr;  // 3 x nwords words
  unsigned* q = nullptr;             // 2 x n entries
  unsigned* qe = nullptr;            // 2 x (n + 1) entries
  unsigned* fq = nullptr;            // near-far runs: 2 x n entries
  unsigned long long* fbm = nullptr; // near-far runs: nwords words
  step_stats* stats = nullptr;       // max_iter entries
  step_ctl* ctl = nullptr;
  group_stats* gst = nullptr;        // kTicketGroups entries
  unsigned* word = nullptr;          // one device word for the largest out-degree (no workspace)
  unsigned nwords = 0;
  unsigned max_iter = 0;
  step_ctl* ctl_host = nullptr;      // host memory the control header is copied into (two
                                     // slots when evc is set, else one)
  hipEvent_t evc[2] = {nullptr, nullptr};  // completion of the copy into each slot (pipelining)
  launch_cache* graphs = nullptr;    // null: iteration chunks are launched directly
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // null: the host clock times the run
};

inline run_state state_of(workspace& ws) {
  run_state st;
  st.bm = ws.bm.get_ptr();
  st.q = ws.q.get_ptr();
  st.qe = ws.qe.get_ptr();
  st.fq = ws.fq.size() ? ws.fq.get_ptr() : nullptr;
  st.fbm = ws.fbm.size() ? ws.fbm.get_ptr() : nullptr;
  st.stats = ws.stats.get_ptr();
  st.ctl = ws.ctl.get_ptr();
  st.gst = ws.gst.get_ptr();
  st.nwords = ws.nwords;
  st.max_iter = ws.max_iter;
  st.ctl_host = ws.ctl_host.get_ptr();
  st.graphs = &ws.graphs;
  st.ev0 = ws.ev0;
  st.ev1 = ws.ev1;
  st.evc[0] = ws.evc[0];
  st.evc[1] = ws.evc[1];
  return st;
}

inline unsigned small_grid_of(gpu& dev, options const& o) {
  return o.small_grid ? o.small_grid : static_cast<unsigned>(dev.numComputeUnits());
}

inline unsigned tiny_grid_of(gpu& dev, options const& o) {
  return o.tiny_grid ? o.tiny_grid
                     : std::max(1u, static_cast<unsigned>(dev.numComputeUnits()) / 4u);
}

// Replays chunks of `chunk` launches until the run is done.  Every launch
// of one chunk uses the same grid, picked from the state the last finished
// chunk left (ctl.small_next): the full grid; one workgroup per compute
// unit after a small push frontier (entries + out-edges <= kSmallWork);
// a quarter of that after a tiny one (<= kTinyWork).  Small grids cut the
// fixed cost of an iteration on long runs of small frontiers (high-diameter
// graphs).  Within a launch, a push iteration does its work on as many
// workgroups as its queue needs (at least prm.small_grid), and the rest
// only count in.  With options::persist a tiny-grid chunk starts with
// persist_kernel, which runs the iterations that fit one workgroup until
// the frontier outgrows it; the chunk's step launches take the iterations
// after that (or return at once when the run is done).
//
// The first chunk is waited for (short runs end in it).  After that, with
// options::pipeline and the run state's copy events, the next chunk is
// queued before the previous one's control line is read, so the host's
// round trip overlaps device work; its grid comes from the chunk before,
// and a chunk queued after the run is done only returns at once.
template <bool DELTA, bool MIN_FAST, bool NF, class P>
unsigned run_loop(P const& prm, gpu& dev, stream& s, unsigned grid, options const& o,
                  run_state const& st) {
  auto kern = &step_kernel<DELTA, MIN_FAST, NF, P>;
  unsigned const chunk = std::max(1u, o.chunk);
  unsigned g[3];
  g[0] = grid;
  g[1] = std::min(grid, small_grid_of(dev, o));
  g[2] = std::min(g[1], tiny_grid_of(dev, o));
  bool const levels = o.adaptive_grid && g[1] < grid;
  void (*pre)(P) = nullptr;
  if constexpr (DELTA && MIN_FAST) {
    if (levels && prm.persist_edges &&
        (prm.sched.mode == SCHED_PLAIN || prm.sched.mode == SCHED_LEARNED))
      pre = &persist_kernel<NF, P>;
  }
  // Executables per level, taken from the cache (or captured) when a level
  // is first used.  A capture records launches without running them, so it
  // may follow work still in flight on the stream; what it can drop from
  // the cache is older than the three executables this run uses.
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  bool const graphs = o.use_graph && st.graphs;
  bool const pipe = o.pipeline && st.evc[0] && st.evc[1];
  auto issue = [&](unsigned lv, unsigned slot) {
    if (graphs && !exec[lv]) {
      exec[lv] = st.graphs->chunk(kern, g[lv], COMPILE_TIME_DETERMINED_BLOCK_SIZE, chunk, prm,
                                  s.get(), lv == 2 ? pre : nullptr, kPersistBlock);
    }
    if (exec[lv]) {
      KITES_HIP_CHECK(hipGraphLaunch(exec[lv], s.get()));
    } else {
      if (lv == 2 && pre)
        hipLaunchKernelGGL(pre, dim3(1), dim3(kPersistBlock), 0, s.get(), prm);
      for (unsigned k = 0; k < chunk; ++k)
        hipLaunchKernelGGL(kern, dim3(g[lv]), dim3(COMPILE_TIME_DETERMINED_BLOCK_SIZE), 0,
                           s.get(), prm);
      KITES_HIP_CHECK(hipPeekAtLastError());
    }
    KITES_HIP_CHECK(hipMemcpyAsync(st.ctl_host + slot, st.ctl, offsetof(step_ctl, board),
                                   hipMemcpyDeviceToHost, s.get()));
    if (pipe) KITES_HIP_CHECK(hipEventRecord(st.evc[slot], s.get()));
  };
  auto level_of = [&](step_ctl const& c) { return levels ? std::min(c.small_next, 2u) : 0u; };
  unsigned long long const cap = static_cast<unsigned long long>(st.max_iter) + chunk;
  unsigned long long launched = chunk;
  issue(0, 0);
  s.sync();
  if (st.ctl_host[0].done) return static_cast<unsigned>(launched);
  unsigned lv = level_of(st.ctl_host[0]);
  if (!pipe) {
    while (launched <= cap) {
      issue(lv, 0);
      launched += chunk;
      s.sync();
      if (st.ctl_host[0].done) break;
      lv = level_of(st.ctl_host[0]);
    }
    return static_cast<unsigned>(launched);
  }
  // slot: the chunk whose control line is read next; last: the last queued
  unsigned slot = 0, last = 0;
  issue(lv, slot);
  launched += chunk;
  for (;;) {
    bool const more = launched <= cap;
    if (more) {
      last = slot ^ 1u;
      issue(lv, last);
      launched += chunk;
    }
    KITES_HIP_CHECK(hipEventSynchronize(st.evc[slot]));
    step_ctl const& c = st.ctl_host[slot];
    if (c.done || !more) break;
    lv = level_of(c);
    slot ^= 1u;
  }
  s.sync();
  // the caller reads the final control line from the first slot
  if (last != 0) st.ctl_host[0] = st.ctl_host[last];
  return static_cast<unsigned>(launched);
}

// Iteration loop of the recorder: one launch at a time.  Before each, the
// state the iteration reads (values, control block, its input bitmap and
// queue) is saved; every delta variant is run from it and timed, and the
// state is restored.  With record_rollout = 0 a variant's time is its own
// launch; otherwise it is that launch plus the rest of the run under the
// forced rollout variant (time-to-finish), which also charges a variant
// for the frontier it leaves behind.
template <bool DELTA, bool MIN_FAST, bool NF, class P>
unsigned record_loop(P const& prm, gpu& dev, stream& s, unsigned grid, options const& o,
                     workspace& ws) {
  using vT = typename std::remove_reference<decltype(*prm.V)>::type;
  auto kern = &step_kernel<DELTA, MIN_FAST, NF, P>;
  std::size_t const n = prm.n;
  std::size_t const nw = prm.nwords;
  unsigned const rv = o.record_rollout;
  P roll = prm;
  if (rv) {
    roll.sched.mode = SCHED_FORCED;
    roll.sched.forced_variant = rv;
  }
  device_buffer<vT> vsave(n ? n : 1, ws.dev);
  device_buffer<step_ctl> csave(1, ws.dev);
  device_buffer<unsigned long long> bsave(nw, ws.dev);
  device_buffer<unsigned> qsave(n ? n : 1, ws.dev);
  device_buffer<unsigned> qesave(n + 1, ws.dev);
  hipEvent_t e0, e1;
  KITES_HIP_CHECK(hipEventCreate(&e0));
  KITES_HIP_CHECK(hipEventCreate(&e1));
  constexpr unsigned kRollChunk = 8;
  hipGraph_t rgraph = nullptr;
  hipGraphExec_t rexec = nullptr;
  if (rv) {
    KITES_HIP_CHECK(hipStreamBeginCapture(s.get(), hipStreamCaptureModeThreadLocal));
    for (unsigned k = 0; k < kRollChunk; ++k)
      hipLaunchKernelGGL(kern, dim3(grid), dim3(COMPILE_TIME_DETERMINED_BLOCK_SIZE), 0, s.get(),
                         roll);
    KITES_HIP_CHECK(hipStreamEndCapture(s.get(), &rgraph));
    KITES_HIP_CHECK(hipGraphInstantiate(&rexec, rgraph, nullptr, nullptr, 0));
  }
  unsigned launched = 0;
  auto launch = [&](P const& pp) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(COMPILE_TIME_DETERMINED_BLOCK_SIZE), 0, s.get(), pp);
    KITES_HIP_CHECK(hipPeekAtLastError());
  };
  auto timed = [&](auto&& body) {
    KITES_HIP_CHECK(hipEventRecord(e0, s.get()));
    body();
    KITES_HIP_CHECK(hipEventRecord(e1, s.get()));
    KITES_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    KITES_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  };
  auto restore = [&](unsigned iter) {
    KITES_HIP_CHECK(hipMemcpyAsync(prm.V, vsave.get_ptr(), n * sizeof(vT), hipMemcpyDeviceToDevice,
                                   s.get()));
    KITES_HIP_CHECK(hipMemcpyAsync(prm.ctl, csave.get_ptr(), sizeof(step_ctl),
                                   hipMemcpyDeviceToDevice, s.get()));
    KITES_HIP_CHECK(hipMemcpyAsync(prm.bm + (iter % 3) * nw, bsave.get_ptr(), nw * 8,
                                   hipMemcpyDeviceToDevice, s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(prm.bm + ((iter + 1) % 3) * nw, 0, nw * 8, s.get()));
    KITES_HIP_CHECK(hipMemcpyAsync(prm.q + (iter & 1) * n, qsave.get_ptr(), n * 4,
                                   hipMemcpyDeviceToDevice, s.get()));
    KITES_HIP_CHECK(hipMemcpyAsync(prm.qe + (iter & 1) * (n + 1), qesave.get_ptr(), (n + 1) * 4,
                                   hipMemcpyDeviceToDevice, s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(prm.stats + iter, 0, (ws.max_iter - iter) * sizeof(step_stats),
                                   s.get()));
  };
  for (;;) {
    KITES_HIP_CHECK(hipMemcpyAsync(ws.ctl_host.get_ptr(), prm.ctl, offsetof(step_ctl, board),
                                   hipMemcpyDeviceToHost, s.get()));
    s.sync();
    step_ctl const ctl = ws.ctl_host[0];
    if (ctl.done || launched > ws.max_iter) break;
    unsigned const iter = ctl.iter;
    if (ctl.next_variant != V_CLASSIC && ctl.next_variant != V_INIT_LB) {
      variant_sample smp{};
      unsigned count, maxdeg;
      unsigned long long sdeg;
      if (iter == 0) {
        unsigned sent = 0;
        count = ctl.seed_count;
        KITES_HIP_CHECK(hipMemcpy(&sent, prm.qe + count, sizeof(unsigned), hipMemcpyDeviceToHost));
        sdeg = sent;
        maxdeg = count == 1 ? sent : 0u;
      } else {
        step_stats prev;
        KITES_HIP_CHECK(hipMemcpy(&prev, prm.stats + iter - 1, sizeof(step_stats),
                                  hipMemcpyDeviceToHost));
        count = prev.count;
        sdeg = prev.sum_outdeg;
        maxdeg = prev.max_outdeg;
      }
      make_features(smp.feat, prm.n, prm.m, count, sdeg, maxdeg, iter,
                    prm.sched.settled_below != 0);
      smp.iter = iter;
      smp.plain_variant = ctl.next_variant;
      KITES_HIP_CHECK(hipMemcpyAsync(vsave.get_ptr(), prm.V, n * sizeof(vT),
                                     hipMemcpyDeviceToDevice, s.get()));
      KITES_HIP_CHECK(hipMemcpyAsync(csave.get_ptr(), prm.ctl, sizeof(step_ctl),
                                     hipMemcpyDeviceToDevice, s.get()));
      KITES_HIP_CHECK(hipMemcpyAsync(bsave.get_ptr(), prm.bm + (iter % 3) * nw, nw * 8,
                                     hipMemcpyDeviceToDevice, s.get()));
      KITES_HIP_CHECK(hipMemcpyAsync(qsave.get_ptr(), prm.q + (iter & 1) * n, n * 4,
                                     hipMemcpyDeviceToDevice, s.get()));
      KITES_HIP_CHECK(hipMemcpyAsync(qesave.get_ptr(), prm.qe + (iter & 1) * (n + 1), (n + 1) * 4,
                                     hipMemcpyDeviceToDevice, s.get()));
      // Edge-balanced variants first, so that the fastest time so far bounds
      // how often a much slower variant is repeated.
      static constexpr unsigned order[MLP_OUT] = {V_PUSH_LB, V_PULL_LB, V_PUSH_T, V_PULL_T,
                                                  V_PUSH_W, V_PULL_W, V_PUSH_B, V_PULL_B};
      float fastest = 0.f;
      for (unsigned v : order) {
        float best = 0.f;
        for (unsigned r = 0; r < std::max(1u, o.record_repeats); ++r) {
          if (r > 0 && best > 4.f * fastest) break;
          restore(iter);
          KITES_HIP_CHECK(hipMemcpyAsync(reinterpret_cast<char*>(prm.ctl) + offsetof(step_ctl, next_variant),
                                         &v, sizeof(unsigned), hipMemcpyHostToDevice, s.get()));
          float ms = timed([&] { launch(rv ? roll : prm); });
          // rollout: replay chunks of forced launches until the run ends;
          // each chunk is timed on its own so host checks add nothing
          for (unsigned k = 0; rv && k <= ws.max_iter / kRollChunk + 1; ++k) {
            KITES_HIP_CHECK(hipMemcpyAsync(ws.ctl_host.get_ptr(), prm.ctl, offsetof(step_ctl, board),
                                           hipMemcpyDeviceToHost, s.get()));
            s.sync();
            if (ws.ctl_host[0].done) break;
            ms += timed([&] { KITES_HIP_CHECK(hipGraphLaunch(rexec, s.get())); });
          }
          best = (r == 0 || ms < best) ? ms : best;
        }
        smp.ms[v - 1] = best;
        fastest = (v == order[0] || best < fastest) ? best : fastest;
      }
      restore(iter);
      s.sync();
      o.record->push_back(smp);
    }
    launch(prm);
    ++launched;
  }
  if (rexec) (void)hipGraphExecDestroy(rexec);
  if (rgraph) (void)hipGraphDestroy(rgraph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return launched;
}

}  // namespace detail

namespace detail {

// The ordered mode's loop: per iteration, copy the values, clear the next
// frontier and the flag, run ordered_kernel, read the flag; the values and
// the frontiers alternate between two buffers.
template <class vT, class eT, class vIdxT, class FI, class FN, class FR, class FU>
void run_ordered(graph_csr_device<vT, eT, vIdxT>& grph, gpu& dev, stream& s, workspace& ws,
                 FI fi, FN fn, FR fr, FU fu, options const& o) {
  std::size_t const n = grph.get_num_vertices();
  int const did = grph.getResidingDeviceID();
  device_buffer<vT> vb(n ? n : 1, did);
  unsigned long long* const bm0 = ws.bm.get_ptr();
  unsigned long long* const bm1 = bm0 + ws.nwords;
  std::size_t const wbytes = static_cast<std::size_t>(ws.nwords) * sizeof(unsigned long long);
  KITES_HIP_CHECK(hipEventRecord(ws.ev0, s.get()));
  KITES_HIP_CHECK(hipMemsetAsync(bm0, 0, 2 * wbytes, s.get()));
  // the initial frontier is consumed, as in the other schedules
  std::size_t const fb = std::min<std::size_t>(grph.vBitmapEven.sizeInBytes(), wbytes);
  if (fb) {
    KITES_HIP_CHECK(hipMemcpyAsync(bm0, grph.vBitmapEven.get_ptr(), fb, hipMemcpyDeviceToDevice,
                                   s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(grph.vBitmapEven.get_ptr(), 0, grph.vBitmapEven.sizeInBytes(),
                                   s.get()));
  }
  using P = ordered_params<vT, eT, FI, FN, FR, FU>;
  param_block<P> pb(fi, fn, fr, fu);
  P& prm = pb.get();
  prm.n = static_cast<unsigned>(n);
  prm.R = reinterpret_cast<const unsigned*>(grph.R.get_ptr());
  prm.C = reinterpret_cast<const unsigned*>(grph.C.get_ptr());
  prm.E = grph.E.get_ptr();
  prm.oR = reinterpret_cast<const unsigned*>(grph.oR.get_ptr());
  prm.oC = reinterpret_cast<const unsigned*>(grph.oC.get_ptr());
  prm.nwords = ws.nwords;
  prm.flag = grph.updateFlag.get_ptr();
  vT* vin = grph.V.get_ptr();
  vT* vout = vb.get_ptr();
  unsigned long long* bin = bm0;
  unsigned long long* bout = bm1;
  unsigned const grid = static_cast<unsigned>(dev.numComputeUnits()) * 8u;
  unsigned const max_iter = std::max(1u, o.max_iterations);
  unsigned iters = 0;
  bool done = false;
  while (iters < max_iter) {
    if (n)
      KITES_HIP_CHECK(hipMemcpyAsync(vout, vin, n * sizeof(vT), hipMemcpyDeviceToDevice, s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(bout, 0, wbytes, s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(prm.flag, 0, sizeof(unsigned), s.get()));
    prm.Vin = vin;
    prm.Vout = vout;
    prm.in_bm = bin;
    prm.out_bm = bout;
    auto const kern = &ordered_kernel<256, P>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s.get(), prm);
    KITES_HIP_CHECK(hipPeekAtLastError());
    KITES_HIP_CHECK(hipMemcpyAsync(grph.updateFlagHost.get_ptr(), prm.flag, sizeof(unsigned),
                                   hipMemcpyDeviceToHost, s.get()));
    s.sync();
    ++iters;
    std::swap(vin, vout);
    std::swap(bin, bout);
    if (grph.updateFlagHost[0] == 0) {
      done = true;
      break;
    }
  }
  if (vin != grph.V.get_ptr() && n)
    KITES_HIP_CHECK(hipMemcpyAsync(grph.V.get_ptr(), vin, n * sizeof(vT), hipMemcpyDeviceToDevice,
                                   s.get()));
  KITES_HIP_CHECK(hipEventRecord(ws.ev1, s.get()));
  s.sync();
  float ms = 0.f;
  KITES_HIP_CHECK(hipEventElapsedTime(&ms, ws.ev0, ws.ev1));
  if (!done) throw std::runtime_error("kites::process: iteration limit reached before convergence");
  if (o.verbose) {
    std::cout << "Processing finished in " << ms << " (ms).\n";
    std::cout << "Performed " << iters << " iterations in total.\n";
  }
  if (o.report) {
    o.report->iterations = iters;
    o.report->ms = ms;
    o.report->delta = false;
    o.report->min_fast = false;
  }
}

}  // namespace detail

namespace detail {

// Largest out-degree of the graph whose out-offsets are oR (nv vertices),
// reduced into the device word `w` and read back (no allocation).
inline unsigned max_outdeg_into(unsigned* w, const unsigned* oR, std::size_t nv, stream& s) {
  hipLaunchKernelGGL(zero_words_kernel, dim3(1), dim3(64), 0, s.get(), w, 1u);
  if (nv) {
    unsigned const blocks = static_cast<unsigned>(std::min<std::size_t>((nv + 255) / 256, 4096));
    hipLaunchKernelGGL(max_outdeg_kernel, dim3(blocks), dim3(256), 0, s.get(), oR,
                       static_cast<unsigned>(nv), w);
  }
  KITES_HIP_CHECK(hipPeekAtLastError());
  unsigned md = 0;
  KITES_HIP_CHECK(hipMemcpyAsync(&md, w, sizeof(unsigned), hipMemcpyDeviceToHost, s.get()));
  s.sync();
  return md;
}

// The fixpoint driver over a run state `st` (see run_state).  `ws` is the
// graph's workspace, or null when the state is caller-owned: then nothing
// is allocated (no recorder, no split pull, no hipGraph capture, no events).
template <class vT, class eT, class vIdxT, class funcInitT, class funcCompNbrT, class funcRedT,
          class funcUpdT>
void run_engine(graph_csr_device<vT, eT, vIdxT>& grph, gpu& dev, funcInitT funcInit,
                funcCompNbrT funcCompNbr, funcRedT funcRed, funcUpdT funcUpd, options const& o,
                run_state& st, workspace* ws) {
  stream& s = dev.getSsEven();
  std::size_t const n = grph.get_num_vertices();
  unsigned const vpg = grph.getVertexPerGroup();
  if (vpg == 0 || (vpg & (vpg - 1)) != 0)
    throw std::runtime_error("kites::process: vertex-per-group must be a power of two");
  unsigned vpg_shift = 0;
  while ((1u << vpg_shift) < vpg) ++vpg_shift;
  if (vpg > 1 && (grph.dir_R.size() == 0 || grph.gProp != graph_property::directed))
    throw std::runtime_error("kites::process: vertex-per-group > 1 needs a directed graph");
  if (!ws && (o.ordered || o.split_pull || o.record))
    throw std::invalid_argument(
        "kites::process: the ordered mode, the split pull and the recorder need the graph's "
        "workspace");
  if (o.ordered) {
    if (vpg != 1) throw std::runtime_error("kites::process: the ordered mode needs vertex-per-group 1");
    run_ordered(grph, dev, s, *ws, funcInit, funcCompNbr, funcRed, funcUpd, o);
    return;
  }
  probe_result const pr = probe_functors<vT, eT>(funcInit, funcCompNbr, funcRed, funcUpd);
  schedule sch = o.sched;
  if (sch == schedule::automatic)
    sch = (pr.delta && vpg == 1) ? schedule::plain : schedule::classic;
  if (sch != schedule::classic && !pr.delta)
    throw std::runtime_error(
        "kites::process: the requested schedule needs a min-semilattice functor set");
  if (sch != schedule::classic && vpg != 1)
    throw std::runtime_error("kites::process: delta schedules need vertex-per-group 1");
  if (sch == schedule::learned && !o.mlp_weights)
    throw std::runtime_error("kites::process: learned schedule without model weights");
  if (o.seed != options::seeding::none && sch == schedule::classic)
    throw std::runtime_error("kites::process: a seeded start needs a delta schedule");
  if (o.seed == options::seeding::source && o.seed_source >= n)
    throw std::runtime_error("kites::process: seed source out of range");
  if (o.seed == options::seeding::list && (o.seed_count > n || (o.seed_count && !o.seed_list)))
    throw std::runtime_error("kites::process: bad seed list");

  using P = step_params<vT, eT, funcInitT, funcCompNbrT, funcRedT, funcUpdT>;
  step_core<vT, eT> prm;
  zero_params(prm);
  prm.n = static_cast<unsigned>(n);
  prm.m = grph.get_num_edges();
  prm.R = grph.R.get_ptr();
  prm.C = grph.C.get_ptr();
  prm.E = grph.E.get_ptr();
  prm.oR = grph.oR.get_ptr();
  prm.oC = grph.oC.get_ptr();
  prm.oE = grph.oE.get_ptr();
  prm.nzi = grph.nzi.get_ptr();
  prm.nzR = grph.nzR.get_ptr();
  prm.outbits = grph.outbits.get_ptr();
  prm.nnz = grph.nnz;
  prm.gR = grph.dir_R.get_ptr();
  prm.gC = grph.dir_C.get_ptr();
  prm.ngroups = grph.dir_R.size() ? static_cast<unsigned>(grph.dir_R.size() - 1) : 0u;
  prm.V = grph.V.get_ptr();
  prm.vpg_shift = vpg_shift;
  prm.vpg = vpg;
  prm.bm = st.bm;
  prm.q = st.q;
  prm.qe = st.qe;
  prm.fq = o.delta ? st.fq : nullptr;
  prm.fbm = o.delta ? st.fbm : nullptr;
  if (o.delta && (!st.fq || !st.fbm))
    throw std::invalid_argument("kites::process: the near-far run state has no far pile");
  prm.nwords = st.nwords;
  prm.stats = st.stats;
  prm.ctl = st.ctl;
  prm.update_flag = grph.updateFlag.get_ptr();
  prm.sched.mode = sch == schedule::classic ? SCHED_CLASSIC
                   : sch == schedule::plain ? SCHED_PLAIN
                   : sch == schedule::forced ? SCHED_FORCED
                                             : SCHED_LEARNED;
  prm.sched.fixed_gran = static_cast<unsigned>(o.gran);
  prm.sched.forced_variant = o.forced_variant;
  prm.sched.alpha = o.alpha;
  prm.sched.margin = o.margin;
  prm.sched.delta = o.delta;
  prm.sched.settled_below = o.settled_below;
  prm.sched.mlp = o.mlp_weights;
  prm.collect = (o.keep_stats || o.record) ? 1u : 0u;
  prm.gst = st.gst;
  prm.board = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(st.ctl) +
                                          offsetof(step_ctl, board));
  prm.small_grid = o.adaptive_grid ? small_grid_of(dev, o) : 0u;
  prm.persist_edges = o.persist ? std::min(o.persist_edges, detail::kPersistEdges) : 0u;
  if (sch == schedule::learned)
    prm.max_outdeg = ws ? ws->max_outdeg_of(grph.oR.get_ptr(), n, s)
                        : max_outdeg_into(st.word, grph.oR.get_ptr(), n, s);
  if (o.split_pull && sch != schedule::classic) {
    ensure_split(grph, s);
    prm.xsplit = 1u;
    prm.xC = reinterpret_cast<const unsigned*>(grph.xC.get_ptr());
    prm.xE = grph.xE.get_ptr();
    prm.xnzi = reinterpret_cast<const unsigned*>(grph.xnzi.get_ptr());
    prm.xnzR = reinterpret_cast<const unsigned*>(grph.xnzR.get_ptr());
    for (unsigned b = 0; b <= kSplitParts; ++b) prm.xseg[b] = grph.xseg[b];
  }
  param_block<P> const kpb(prm, funcInit, funcCompNbr, funcRed, funcUpd);
  P const& kp = kpb.get();

  auto const t0 = std::chrono::steady_clock::now();
  if (st.ev0) KITES_HIP_CHECK(hipEventRecord(st.ev0, s.get()));
  // Reset the run state (all three bitmap slots included), then load the
  // initial frontier (consumed, as the reference leaves both of its bitmaps
  // empty after a run).
  hipLaunchKernelGGL(reset_kernel, dim3(256), dim3(256), 0, s.get(), prm.stats, st.max_iter,
                     prm.bm, st.nwords, prm.ctl,
                     static_cast<unsigned>(sch == schedule::classic ? V_CLASSIC : V_INIT_LB),
                     std::min(st.max_iter, std::max(2u, o.max_iterations)), prm.update_flag,
                     prm.gst, prm.board, static_cast<unsigned long long>(o.delta), prm.fbm);
  KITES_HIP_CHECK(hipPeekAtLastError());
  std::size_t const fb = std::min<std::size_t>(grph.vBitmapEven.sizeInBytes(),
                                               st.nwords * sizeof(unsigned long long));
  if (fb) {
    KITES_HIP_CHECK(hipMemcpyAsync(prm.bm, grph.vBitmapEven.get_ptr(), fb,
                                   hipMemcpyDeviceToDevice, s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(grph.vBitmapEven.get_ptr(), 0, grph.vBitmapEven.sizeInBytes(),
                                   s.get()));
  }

  if (o.seed != options::seeding::none) {
    unsigned const mode = o.seed == options::seeding::source ? SEED_SOURCE
                          : o.seed == options::seeding::all  ? SEED_ALL
                                                              : SEED_LIST;
    if (fb)
      KITES_HIP_CHECK(hipMemsetAsync(prm.bm, 0, st.nwords * sizeof(unsigned long long), s.get()));
    hipLaunchKernelGGL(seed_kernel<P>, dim3(mode == SEED_ALL ? 1024 : 1),
                       dim3(COMPILE_TIME_DETERMINED_BLOCK_SIZE), 0, s.get(), kp, mode,
                       o.seed_source, o.seed_list, o.seed_count);
    KITES_HIP_CHECK(hipPeekAtLastError());
  }

  unsigned grid = 0;
  bool const use_delta = (sch != schedule::classic);
  bool used_min_fast = false;
  if (o.record && o.delta)
    throw std::invalid_argument("kites::process: the recorder does not fork near-far runs");
  if (o.record && !use_delta)
    throw std::runtime_error("kites::process: the recorder needs a delta schedule");
  if (use_delta) {
    if constexpr (min_ops<vT>::ok) {
      if (pr.min_fast) {
        if constexpr (std::is_unsigned<vT>::value) {
          if (o.delta) {
            // near-far push: its own kernel, so runs without it keep the
            // smaller register footprint
            grid = step_grid<true, true, true, P>(dev, o.grid);
            if (o.record)
              record_loop<true, true, true>(kp, dev, s, grid, o, *ws);
            else
              run_loop<true, true, true>(kp, dev, s, grid, o, st);
            used_min_fast = true;
          }
        }
        if (!used_min_fast) {
          grid = step_grid<true, true, false, P>(dev, o.grid);
          if (o.record)
            record_loop<true, true, false>(kp, dev, s, grid, o, *ws);
          else
            run_loop<true, true, false>(kp, dev, s, grid, o, st);
          used_min_fast = true;
        }
      }
    }
    if (!used_min_fast) {
      if constexpr (sizeof(vT) == 4 || sizeof(vT) == 8) {
        grid = step_grid<true, false, false, P>(dev, o.grid);
        if (o.record)
          record_loop<true, false, false>(kp, dev, s, grid, o, *ws);
        else
          run_loop<true, false, false>(kp, dev, s, grid, o, st);
      }
    }
  } else {
    grid = step_grid<false, false, false, P>(dev, o.grid);
    run_loop<false, false, false>(kp, dev, s, grid, o, st);
  }
  if (st.ev1) KITES_HIP_CHECK(hipEventRecord(st.ev1, s.get()));
  s.sync();
  float ms = 0.f;
  if (st.ev0 && st.ev1)
    KITES_HIP_CHECK(hipEventElapsedTime(&ms, st.ev0, st.ev1));
  else
    ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  step_ctl const ctl = *st.ctl_host;
  if (!ctl.done || ctl.overflow)
    throw std::runtime_error("kites::process: iteration limit reached before convergence");
  if (o.verbose) {
    std::cout << "Processing finished in " << ms << " (ms).\n";
    std::cout << "Performed " << ctl.iter << " iterations in total.\n";
  }
  if (o.report) {
    o.report->iterations = ctl.iter;
    o.report->ms = ms;
    o.report->delta = use_delta;
    o.report->min_fast = used_min_fast;
    o.report->graph_builds = st.graphs ? st.graphs->builds() : 0u;
    if (o.keep_stats) {
      o.report->stats.resize(ctl.iter);
      if (ctl.iter)
        KITES_HIP_CHECK(hipMemcpy(o.report->stats.data(), prm.stats,
                                  ctl.iter * sizeof(step_stats), hipMemcpyDeviceToHost));
    }
  }
}

}  // namespace detail

// Runs the fixpoint with explicit options (extension of the reference API).
template <launch launchMode, class vT, class eT, class vIdxT, class funcInitT,
          class funcCompNbrT, class funcRedT, class funcUpdT>
void process(graph_csr_device<vT, eT, vIdxT>& grph, gpu& dev, funcInitT funcInit,
             funcCompNbrT funcCompNbr, funcRedT funcRed, funcUpdT funcUpd, options const& o) {
  using namespace detail;
  dev.setAsActive();
  std::size_t const n = grph.get_num_vertices();
  if (o.verbose) {
    auto const mu = dev.queryMemUsage();
    std::cout << mu.first << " MB is consumed out of the total " << mu.second
              << " MB global memory.\n";
    std::cout << "Launching iterative procedure ...\n";
  }
  if (!grph.ws) grph.ws.reset(new workspace());
  workspace& ws = *grph.ws;
  ws.ensure(n, std::max(2u, o.max_iterations), grph.getResidingDeviceID());
  if (o.delta) ws.ensure_far();
  run_state st = state_of(ws);
  run_engine(grph, dev, funcInit, funcCompNbr, funcRed, funcUpd, o, st, &ws);
}

// Runs the fixpoint on a run state the caller provides (device memory for
// the bitmap ring, queues, statistics, control block and group statistics,
// sized for the graph's n and st.max_iter), allocating nothing: no
// workspace, no hipGraph capture, no events (the host clock times the run).
// Chunks of o.chunk iterations are launched directly.  The ordered mode, the
// split pull, the recorder and per-iteration statistics are refused.
template <launch launchMode, class vT, class eT, class vIdxT, class funcInitT,
          class funcCompNbrT, class funcRedT, class funcUpdT>
void process_noalloc(graph_csr_device<vT, eT, vIdxT>& grph, gpu& dev, funcInitT funcInit,
                     funcCompNbrT funcCompNbr, funcRedT funcRed, funcUpdT funcUpd,
                     options const& o, detail::run_state& st) {
  if (o.keep_stats || o.verbose)
    throw std::invalid_argument("kites::process_noalloc: statistics and verbose output allocate");
  std::size_t const n = grph.get_num_vertices();
  if (!st.bm || !st.q || !st.qe || !st.stats || !st.ctl || !st.gst || !st.word || !st.ctl_host)
    throw std::invalid_argument("kites::process_noalloc: incomplete run state");
  if (st.nwords < (n + 63) / 64 + 1)
    throw std::invalid_argument("kites::process_noalloc: run state too small for the graph");
  if (std::max(2u, o.max_iterations) > st.max_iter)
    throw std::invalid_argument("kites::process_noalloc: iteration limit above the run state's");
  dev.setAsActive();
  options oo = o;
  oo.use_graph = false;
  st.graphs = nullptr;
  st.ev0 = st.ev1 = nullptr;
  detail::run_engine(grph, dev, funcInit, funcCompNbr, funcRed, funcUpd, oo, st, nullptr);
}

// Drop-in: process<launch::sync>(devGrph, dev, initF, nbrCompF, redF, updF)
// (compute.cuh:29-38).  launch::async is accepted and runs synchronously, as
// in the reference (compute.cuh:408).
template <launch launchMode, class vT, class eT, class vIdxT, class funcInitT,
          class funcCompNbrT, class funcRedT, class funcUpdT>
void process(graph_csr_device<vT, eT, vIdxT>& grph, gpu& dev, funcInitT funcInit,
             funcCompNbrT funcCompNbr, funcRedT funcRed, funcUpdT funcUpd) {
  options o;
  o.verbose = true;
  process<launchMode>(grph, dev, funcInit, funcCompNbr, funcRed, funcUpd, o);
}

// Drop-in: process on a host graph (compute.cuh:397-411): upload, run, copy
// the vertex values back into the host graph.
template <launch launchMode, class vT, class eT, class vIdxT, class funcInitT,
          class funcCompNbrT, class funcRedT, class funcUpdT>
void process(graph_csr<vT, eT, vIdxT>& grph, gpu& dev, funcInitT funcInit,
             funcCompNbrT funcCompNbr, funcRedT funcRed, funcUpdT funcUpd) {
  auto d = make_graph_csr_for_device(grph, dev);
  process<launch::sync>(d, dev, funcInit, funcCompNbr, funcRed, funcUpd);
  copyMem(grph.V, d.V);
}

}  // namespace kites

#endif  // KITES_ENGINE_HPP_
