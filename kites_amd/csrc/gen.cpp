// gen.cpp -- seeded synthetic graphs and host CSR construction.
//
// Every graph is a pure function of its parameters: edge k of a generator
// draws its random numbers from a splitmix64 stream started at a value that
// depends only on (seed, k), so generation is parallel and reproducible.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <unordered_set>

#include "hgraph.hpp"

namespace kites_amd {

namespace {

inline uint64_t stream_start(uint64_t seed, uint64_t k) {
  return seed * 0x100000001B3ull + k * 0xD1B54A32D192ED03ull + 0x5851F42D4C957F2Dull;
}

inline double unit(uint64_t& st) {
  return static_cast<double>(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
}

inline uint32_t weight(uint64_t& st, uint32_t wmin, uint32_t wmax) {
  uint64_t const span = static_cast<uint64_t>(wmax) - wmin + 1;
  return wmin + static_cast<uint32_t>(splitmix64(st) % span);
}

// Edge k of an R-MAT graph (Chakrabarti, Zhan, Faloutsos, SDM 2004): one
// quadrant choice per bit level, most significant bit first; quadrant
// probabilities a (src 0, dst 0), b (0, 1), c (1, 0), d (1, 1).
inline void rmat_edge(uint64_t seed, uint64_t k, uint32_t scale, double a, double b, double c,
                      uint32_t wmin, uint32_t wmax, uint32_t& s, uint32_t& d, uint32_t& w) {
  uint64_t st = stream_start(seed, k);
  s = 0;
  d = 0;
  double const ab = a + b, abc = a + b + c;
  for (uint32_t lvl = 0; lvl < scale; ++lvl) {
    double const u = unit(st);
    uint32_t sb, db;
    if (u < a) {
      sb = 0; db = 0;
    } else if (u < ab) {
      sb = 0; db = 1;
    } else if (u < abc) {
      sb = 1; db = 0;
    } else {
      sb = 1; db = 1;
    }
    s = (s << 1) | sb;
    d = (d << 1) | db;
  }
  w = weight(st, wmin, wmax);
}

int nthreads(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

}  // namespace

uint32_t pad_vertices(uint32_t n) { return (n % 128) ? (n / 128 + 1) * 128 : n; }

void transpose(hgraph& g, int threads) {
  int const nt = nthreads(threads);
  uint32_t const n = g.n;
  uint64_t const m = g.m;
  std::vector<std::atomic<uint32_t>> cnt(n + 1);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i <= (int64_t)n; ++i) cnt[i].store(0, std::memory_order_relaxed);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t e = 0; e < (int64_t)m; ++e) cnt[g.C[e]].fetch_add(1, std::memory_order_relaxed);
  g.oR.assign(n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) g.oR[v + 1] = g.oR[v] + cnt[v].load(std::memory_order_relaxed);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) cnt[i].store(g.oR[i], std::memory_order_relaxed);
  g.oC.resize(m);
  g.oE.resize(m);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; ++v)
    for (uint32_t e = g.R[v]; e < g.R[v + 1]; ++e) {
      uint32_t const k = cnt[g.C[e]].fetch_add(1, std::memory_order_relaxed);
      g.oC[k] = static_cast<uint32_t>(v);
      g.oE[k] = g.E[e];
    }
  // Destinations ascending within each out-list (the order the reference's
  // dir_C has for vertex-per-group 1).
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)n; ++u) {
    uint32_t const b = g.oR[u], e = g.oR[u + 1];
    if (e - b < 2) continue;
    std::vector<std::pair<uint32_t, uint32_t>> tmp(e - b);
    for (uint32_t i = b; i < e; ++i) tmp[i - b] = {g.oC[i], g.oE[i]};
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](auto const& x, auto const& y) { return x.first < y.first; });
    for (uint32_t i = b; i < e; ++i) {
      g.oC[i] = tmp[i - b].first;
      g.oE[i] = tmp[i - b].second;
    }
  }
}

void build_csr(uint32_t n, std::vector<uint32_t>& src, std::vector<uint32_t>& dst,
               std::vector<uint32_t>& w, bool dedup, int threads, hgraph& g) {
  int const nt = nthreads(threads);
  uint64_t const md = src.size();
  std::vector<std::atomic<uint64_t>> cnt(static_cast<std::size_t>(n) + 1);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i <= (int64_t)n; ++i) cnt[i].store(0, std::memory_order_relaxed);
  for (uint64_t k = 0; k < md; ++k)
    if (dst[k] >= n || src[k] >= n) throw std::runtime_error("edge endpoint out of range");
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t k = 0; k < (int64_t)md; ++k) cnt[dst[k]].fetch_add(1, std::memory_order_relaxed);
  std::vector<uint64_t> off(static_cast<std::size_t>(n) + 1, 0);
  for (uint32_t v = 0; v < n; ++v) off[v + 1] = off[v] + cnt[v].load(std::memory_order_relaxed);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) cnt[i].store(off[i], std::memory_order_relaxed);
  // key = (src << 32) | draw index; sorting a row by key orders it by source
  // and keeps the first-drawn copy of a duplicate first.
  std::vector<uint64_t> key(md);
  std::vector<uint32_t> wt(md);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t k = 0; k < (int64_t)md; ++k) {
    uint64_t const pos = cnt[dst[k]].fetch_add(1, std::memory_order_relaxed);
    key[pos] = (static_cast<uint64_t>(src[k]) << 32) | static_cast<uint64_t>(k & 0xffffffffull);
    wt[pos] = w[k];
  }
  std::vector<uint32_t>().swap(src);
  std::vector<uint32_t>().swap(dst);
  std::vector<uint32_t>().swap(w);
  std::vector<uint32_t> keep(n, 0);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
  for (int64_t v = 0; v < (int64_t)n; ++v) {
    uint64_t const b = off[v], e = off[v + 1];
    if (e - b > 1) {
      std::vector<std::pair<uint64_t, uint32_t>> tmp(e - b);
      for (uint64_t i = b; i < e; ++i) tmp[i - b] = {key[i], wt[i]};
      std::sort(tmp.begin(), tmp.end(),
                [](auto const& x, auto const& y) { return x.first < y.first; });
      for (uint64_t i = b; i < e; ++i) {
        key[i] = tmp[i - b].first;
        wt[i] = tmp[i - b].second;
      }
    }
    uint32_t kept = 0;
    uint64_t prev_src = ~0ull;
    for (uint64_t i = b; i < e; ++i) {
      uint64_t const s = key[i] >> 32;
      if (dedup && (s == static_cast<uint64_t>(v) || s == prev_src)) continue;
      prev_src = s;
      key[b + kept] = key[i];
      wt[b + kept] = wt[i];
      ++kept;
    }
    keep[v] = kept;
  }
  g.n = n;
  g.R.assign(static_cast<std::size_t>(n) + 1, 0);
  for (uint32_t v = 0; v < n; ++v) {
    uint64_t const nx = static_cast<uint64_t>(g.R[v]) + keep[v];
    if (nx > 0xffffffffull) throw std::runtime_error("graph exceeds 2^32 edges");
    g.R[v + 1] = static_cast<uint32_t>(nx);
  }
  g.m = g.R[n];
  g.C.resize(g.m);
  g.E.resize(g.m);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4096)
  for (int64_t v = 0; v < (int64_t)n; ++v) {
    uint64_t const b = off[v];
    for (uint32_t i = 0; i < keep[v]; ++i) {
      g.C[g.R[v] + i] = static_cast<uint32_t>(key[b + i] >> 32);
      g.E[g.R[v] + i] = wt[b + i];
    }
  }
  transpose(g, threads);
}

void gen_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed, double a, double b, double c,
              uint32_t wmin, uint32_t wmax, uint32_t n_limit, uint64_t exact_edges, int threads,
              hgraph& g) {
  if (scale == 0 || scale > 30) throw std::runtime_error("R-MAT scale out of range");
  if (wmin > wmax) throw std::runtime_error("weight range empty");
  uint32_t const nfull = 1u << scale;
  uint32_t const nreal = (n_limit && n_limit < nfull) ? n_limit : nfull;
  std::vector<uint32_t> src, dst, w;
  if (exact_edges) {
    // Draw in order, rejecting out-of-range endpoints, self-loops and
    // duplicates, until exactly `exact_edges` distinct edges exist.
    std::unordered_set<uint64_t> seen;
    seen.reserve(exact_edges * 2);
    uint64_t k = 0;
    uint64_t const cap = exact_edges * 1000 + 1000000;
    while (src.size() < exact_edges) {
      if (k > cap) throw std::runtime_error("R-MAT: cannot reach the requested edge count");
      uint32_t s, d, ww;
      rmat_edge(seed, k++, scale, a, b, c, wmin, wmax, s, d, ww);
      if (s >= nreal || d >= nreal || s == d) continue;
      if (!seen.insert((static_cast<uint64_t>(s) << 32) | d).second) continue;
      src.push_back(s);
      dst.push_back(d);
      w.push_back(ww);
    }
  } else {
    uint64_t const draws = static_cast<uint64_t>(edge_factor) << scale;
    src.resize(draws);
    dst.resize(draws);
    w.resize(draws);
#pragma omp parallel for num_threads(nthreads(threads)) schedule(static)
    for (int64_t k = 0; k < (int64_t)draws; ++k) {
      uint32_t s, d, ww;
      rmat_edge(seed, static_cast<uint64_t>(k), scale, a, b, c, wmin, wmax, s, d, ww);
      if (s >= nreal || d >= nreal) s = d = 0;   // dropped as a self-loop
      src[k] = s;
      dst[k] = d;
      w[k] = ww;
    }
  }
  uint32_t const n = pad_vertices(nreal);
  build_csr(n, src, dst, w, true, threads, g);
  g.n_real = nreal;
}

void gen_er(uint32_t n, uint64_t draws, uint64_t seed, uint32_t wmin, uint32_t wmax, int threads,
            hgraph& g) {
  if (n < 2) throw std::runtime_error("ER graph needs at least 2 vertices");
  std::vector<uint32_t> src(draws), dst(draws), w(draws);
#pragma omp parallel for num_threads(nthreads(threads)) schedule(static)
  for (int64_t k = 0; k < (int64_t)draws; ++k) {
    uint64_t st = stream_start(seed ^ 0x45524552ull, static_cast<uint64_t>(k));
    src[k] = static_cast<uint32_t>(splitmix64(st) % n);
    dst[k] = static_cast<uint32_t>(splitmix64(st) % n);
    w[k] = weight(st, wmin, wmax);
  }
  build_csr(pad_vertices(n), src, dst, w, true, threads, g);
  g.n_real = n;
}

void gen_grid(uint32_t rows, uint32_t cols, uint64_t seed, uint32_t wmin, uint32_t wmax,
              int threads, hgraph& g) {
  uint64_t const nv = static_cast<uint64_t>(rows) * cols;
  if (nv == 0 || nv > 0xffffff00ull) throw std::runtime_error("grid size out of range");
  std::vector<uint32_t> src, dst, w;
  uint64_t const me = 2ull * rows * (cols - 1) + 2ull * cols * (rows - 1);
  src.reserve(me);
  dst.reserve(me);
  w.reserve(me);
  // Directed edges both ways between 4-neighbours, each with its own
  // weight drawn from the stream of its (source, direction) pair.
  const int dr[4] = {-1, 1, 0, 0}, dc[4] = {0, 0, -1, 1};
  for (uint32_t r = 0; r < rows; ++r)
    for (uint32_t c = 0; c < cols; ++c)
      for (int k = 0; k < 4; ++k) {
        int64_t const r2 = (int64_t)r + dr[k], c2 = (int64_t)c + dc[k];
        if (r2 < 0 || c2 < 0 || r2 >= rows || c2 >= cols) continue;
        uint32_t const s = r * cols + c, d = static_cast<uint32_t>(r2 * cols + c2);
        uint64_t st = stream_start(seed ^ 0x47524944ull, static_cast<uint64_t>(s) * 4 + k);
        src.push_back(s);
        dst.push_back(d);
        w.push_back(weight(st, wmin, wmax));
      }
  build_csr(pad_vertices(static_cast<uint32_t>(nv)), src, dst, w, true, threads, g);
  g.n_real = static_cast<uint32_t>(nv);
}

void symmetrize(const hgraph& in, int threads, hgraph& out) {
  std::vector<uint32_t> src, dst, w;
  src.reserve(2 * in.m);
  dst.reserve(2 * in.m);
  w.reserve(2 * in.m);
  for (uint32_t v = 0; v < in.n; ++v)
    for (uint32_t e = in.R[v]; e < in.R[v + 1]; ++e) {
      src.push_back(in.C[e]);
      dst.push_back(v);
      w.push_back(in.E[e]);
    }
  for (uint32_t v = 0; v < in.n; ++v)
    for (uint32_t e = in.R[v]; e < in.R[v + 1]; ++e) {
      src.push_back(v);
      dst.push_back(in.C[e]);
      w.push_back(in.E[e]);
    }
  build_csr(in.n, src, dst, w, true, threads, out);
  out.n_real = in.n_real;
}

}  // namespace kites_amd

namespace kites_amd {

void gen_rmat_batch(uint32_t count, uint64_t seed, uint32_t scale_lo, uint32_t scale_hi,
                    uint32_t ef_lo, uint32_t ef_hi, uint32_t wmin, uint32_t wmax, int threads,
                    hgraph& g, std::vector<uint32_t>& voff, std::vector<uint32_t>& sources) {
  if (count == 0) throw std::runtime_error("batch of zero graphs");
  if (scale_lo == 0 || scale_lo > scale_hi || scale_hi > 24)
    throw std::runtime_error("batch scale range out of bounds");
  if (ef_lo == 0 || ef_lo > ef_hi) throw std::runtime_error("batch edge-factor range empty");
  std::vector<hgraph> parts(count);
  std::vector<uint32_t> pscale(count), pef(count);
  std::vector<uint64_t> pseed(count);
  uint64_t st = seed ^ 0x6b697465736261ull;
  for (uint32_t i = 0; i < count; ++i) {
    pscale[i] = scale_lo + static_cast<uint32_t>(splitmix64(st) % (scale_hi - scale_lo + 1));
    pef[i] = ef_lo + static_cast<uint32_t>(splitmix64(st) % (ef_hi - ef_lo + 1));
    pseed[i] = splitmix64(st);
  }
  // Graphs in parallel, each generated single-threaded (deterministic either way).
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)count; ++i)
    gen_rmat(pscale[i], pef[i], pseed[i], 0.57, 0.19, 0.19, wmin, wmax, 0, 0, 1, parts[i]);
  voff.assign(count + 1, 0);
  uint64_t ntot = 0, mtot = 0;
  for (uint32_t i = 0; i < count; ++i) {
    voff[i] = static_cast<uint32_t>(ntot);
    ntot += parts[i].n;
    mtot += parts[i].m;
  }
  if (ntot >= (1ull << 32) - 1 || mtot >= (1ull << 32))
    throw std::runtime_error("batch too large for 32-bit vertex ids / edge offsets");
  voff[count] = static_cast<uint32_t>(ntot);
  g = hgraph();
  g.n = g.n_real = static_cast<uint32_t>(ntot);
  g.m = mtot;
  g.R.resize(ntot + 1);
  g.oR.resize(ntot + 1);
  g.C.resize(mtot);
  g.E.resize(mtot);
  g.oC.resize(mtot);
  g.oE.resize(mtot);
  std::vector<uint64_t> eoff(count + 1, 0);
  for (uint32_t i = 0; i < count; ++i) eoff[i + 1] = eoff[i] + parts[i].m;
  sources.assign(count, 0);
#pragma omp parallel for num_threads(nthreads(threads)) schedule(dynamic, 1)
  for (int64_t i = 0; i < (int64_t)count; ++i) {
    hgraph const& p = parts[i];
    uint32_t const vo = voff[i];
    uint32_t const eo = static_cast<uint32_t>(eoff[i]);
    for (uint32_t v = 0; v < p.n; ++v) {
      g.R[vo + v] = p.R[v] + eo;
      g.oR[vo + v] = p.oR[v] + eo;
    }
    for (uint64_t e = 0; e < p.m; ++e) {
      g.C[eo + e] = p.C[e] + vo;
      g.E[eo + e] = p.E[e];
      g.oC[eo + e] = p.oC[e] + vo;
      g.oE[eo + e] = p.oE[e];
    }
    // source: max out-degree, lowest id on ties
    uint32_t best = 0, bd = 0;
    for (uint32_t v = 0; v < p.n; ++v) {
      uint32_t const d = p.oR[v + 1] - p.oR[v];
      if (d > bd) {
        bd = d;
        best = v;
      }
    }
    sources[i] = vo + best;
  }
  g.R[ntot] = static_cast<uint32_t>(mtot);
  g.oR[ntot] = static_cast<uint32_t>(mtot);
}

}  // namespace kites_amd
