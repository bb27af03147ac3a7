// host_checks.cpp -- host-only checks of the code that reads untrusted
// input or manages lifetimes on the host, built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/sanitize/Makefile, run by
// tests/test_sanitize.py).  Nothing here touches a GPU.
//
//  * the edge-list parser (kites/graph.hpp, the restatement of the
//    reference's graph_raw.h:58-190) on crafted and random edge lists,
//    checked for the CSR invariants it must produce;
//  * the prepared-graph readers (graph_files.cpp) on every truncation and on
//    random byte corruptions of valid files: each load either throws or
//    yields arrays that satisfy the invariants the kernels rely on;
//  * the graph generators, the edge-list CSR build and the symmetrization;
//  * the launch cache's least-recently-used map (kites/detail/lru.hpp):
//    hits, eviction order, and that every stored value is released exactly
//    once (values own heap memory, so a double or missing release shows up
//    as a sanitizer error or a leak);
//  * the oracle's C functions on small random graphs, cross-checked.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "hgraph.hpp"
#include "kites/detail/lru.hpp"
#include "kites/graph.hpp"
#include "kites_oracle.h"

namespace {

int g_fail = 0;

#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      std::fprintf(stderr, "CHECK failed %s:%d: %s | ", __FILE__, __LINE__, #cond); \
      std::fprintf(stderr, __VA_ARGS__);                   \
      std::fprintf(stderr, "\n");                          \
      ++g_fail;                                            \
    }                                                      \
  } while (0)

using kites_amd::hgraph;

// CSR invariants every host graph must satisfy before upload.
bool csr_ok(uint32_t n, uint64_t m, const std::vector<uint32_t>& R, const std::vector<uint32_t>& C) {
  if (R.size() != static_cast<std::size_t>(n) + 1 || C.size() != m) return false;
  if (R[0] != 0 || R[n] != m) return false;
  for (uint32_t v = 0; v < n; ++v)
    if (R[v] > R[v + 1]) return false;
  for (uint32_t c : C)
    if (c >= n) return false;
  return true;
}

bool hgraph_ok(const hgraph& g) {
  if (!csr_ok(g.n, g.m, g.R, g.C) || g.E.size() != g.m) return false;
  if (!g.oR.empty() && (!csr_ok(g.n, g.m, g.oR, g.oC) || g.oE.size() != g.m)) return false;
  if (!g.V0.empty() && g.V0.size() != g.n) return false;
  if (!g.dir_R.empty()) {
    uint64_t const groups = (static_cast<uint64_t>(g.n) + g.vpg - 1) / g.vpg;
    if (g.dir_R.size() > groups + 1 || g.dir_R[0] != 0 || g.dir_R.back() != g.dir_C.size())
      return false;
    for (std::size_t i = 0; i + 1 < g.dir_R.size(); ++i)
      if (g.dir_R[i] > g.dir_R[i + 1]) return false;
    for (uint32_t x : g.dir_C)
      if (x >= groups) return false;
  }
  return true;
}

// ---- edge-list parser -------------------------------------------------------

struct parse_cfg {
  unsigned vpg;
  kites::input_graph_form form;
  kites::graph_property prop;
  kites::edge_list_expression expr;
};

void parse_and_check(const std::string& text, parse_cfg const& c) {
  std::istringstream in(text);
  // the reference example's callback: weight from the third token (or 1),
  // source value 0 at vertex 30, the rest at 2^30 - 1
  auto cb = [](unsigned argc, char** argv, unsigned s, unsigned d, unsigned& sv, unsigned& dv,
               unsigned& w) {
    w = argc > 0 ? static_cast<unsigned>(std::atoll(argv[0])) : 1u;
    sv = s == 30 ? 0u : 1073741823u;
    dv = d == 30 ? 0u : 1073741823u;
    return s == 30;
  };
  kites::graph_raw<unsigned, unsigned> g(in, cb, c.vpg, c.form, c.prop, c.expr);
  std::size_t const n = g.get_num_vertices();
  CHECK(n % 128 == 0, "n=%zu not padded", n);
  CHECK(g.rawVVec.size() == n && g.rawNbrsVec.size() == n, "sizes");
  std::size_t m = 0;
  for (std::size_t v = 0; v < n; ++v) {
    auto const& l = g.rawNbrsVec[v];
    m += l.size();
    for (std::size_t k = 0; k < l.size(); ++k) {
      CHECK(l[k].nbrIdx < n, "neighbour id %u >= n %zu", l[k].nbrIdx, n);
      if (k) CHECK(!(l[k] < l[k - 1]), "in-list not sorted");
    }
  }
  CHECK(m == g.get_num_edges(), "edge count %zu vs %zu", m, g.get_num_edges());
  std::size_t const groups = (n + g.getVertexPerGroup() - 1) / g.getVertexPerGroup();
  CHECK(g.vBitmap.size() <= groups + 1, "bitmap longer than the groups");
  for (auto const& out : g.rawIVecOutgoing)
    for (unsigned x : out) CHECK(x < groups, "group id %u >= %zu", x, groups);
}

void check_parser() {
  std::vector<parse_cfg> cfgs;
  for (unsigned vpg : {1u, 2u, 4u, 32u})
    for (auto f : {kites::input_graph_form::edge_list_s_d, kites::input_graph_form::edge_list_d_s})
      for (auto p : {kites::graph_property::directed, kites::graph_property::undirected})
        for (auto e : {kites::edge_list_expression::complete, kites::edge_list_expression::incomplete})
          cfgs.push_back({vpg, f, p, e});
  std::vector<std::string> crafted = {
      "",
      "# only a comment\n",
      "1 2\r\n2 3 7\r\n# c\r\n3 1 9 extra tokens here\r\n",
      "0\t1\t5\n1\t\t2\n  3 4\n",
      "5\n6 \n7 8",
      std::string(1000, '9') + " 1\n",                         // atoll saturates nothing: id wraps
      "10 20 " + std::string(400, 'x') + "\n30 40\n",
      "1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 17 18 19 20 21 22 23 24 25 26 27 28 29 30 31 32 "
      "33 34 35 36 37 38 39 40 41 42 43 44 45 46 47 48 49 50 51 52 53 54 55 56 57 58 59 60 61 "
      "62 63 64 65 66 67 68 69 70\n",
      "30 31 2\n31 30 2\n30 30 0\n30 31 1\n",
      "127 0\n128 0\n255 256\n",
      "\n\n\n1 1\n",
      "4+2 3-1\n",
  };
  // the 1000-digit id wraps to a huge vertex id; keep only inputs whose ids
  // stay small enough to allocate
  crafted.erase(crafted.begin() + 5);
  for (auto const& t : crafted)
    for (auto const& c : cfgs) parse_and_check(t, c);
  // random edge lists over a small alphabet (ids below 1000)
  std::mt19937_64 rng(12345);
  const char alpha[] = "0123456789  \t\t\r#x-+";
  for (int it = 0; it < 400; ++it) {
    std::string t;
    int const lines = static_cast<int>(rng() % 40);
    for (int l = 0; l < lines; ++l) {
      int const len = static_cast<int>(rng() % 24);
      int digits = 0;
      for (int k = 0; k < len; ++k) {
        char ch = alpha[rng() % (sizeof(alpha) - 1)];
        if (ch >= '0' && ch <= '9') {
          if (++digits > 3) ch = ' ';
        } else {
          digits = 0;
        }
        if (ch == '-') ch = ' ';   // a negative id wraps to 2^32 - k
        t.push_back(ch);
      }
      t.push_back('\n');
    }
    parse_and_check(t, cfgs[rng() % cfgs.size()]);
  }
}

// ---- prepared-graph files ---------------------------------------------------

hgraph small_graph(uint32_t vpg) {
  hgraph g;
  kites_amd::gen_rmat(7, 4, 99, 0.57, 0.19, 0.19, 1, 9, 0, 0, 2, g);
  g.V0.assign(g.n, 7u);
  g.frontier0.assign((g.n + vpg - 1) / vpg, 0);
  g.frontier0[1] = 1;
  g.vpg = vpg;
  if (vpg > 1) {
    uint32_t const groups = (g.n + vpg - 1) / vpg;
    std::vector<std::vector<uint32_t>> out(groups);
    for (uint32_t v = 0; v < g.n; ++v)
      for (uint32_t e = g.R[v]; e < g.R[v + 1]; ++e) out[g.C[e] / vpg].push_back(v / vpg);
    g.dir_R.assign(1, 0);
    for (auto& l : out) {
      std::sort(l.begin(), l.end());
      l.erase(std::unique(l.begin(), l.end()), l.end());
      g.dir_C.insert(g.dir_C.end(), l.begin(), l.end());
      g.dir_R.push_back(static_cast<uint32_t>(g.dir_C.size()));
    }
  }
  return g;
}

template <class Load>
void corrupt_and_load(const std::string& good, Load load, const char* what, bool every_prefix) {
  // every truncation
  std::size_t const step = every_prefix ? 1 : std::max<std::size_t>(1, good.size() / 97);
  for (std::size_t len = 0; len < good.size(); len += step) {
    std::istringstream in(good.substr(0, len));
    hgraph g;
    try {
      load(in, g);
      CHECK(hgraph_ok(g), "%s: truncation to %zu bytes loaded inconsistent arrays", what, len);
    } catch (std::exception const&) {
    }
  }
  // random corruptions: 1-4 bytes overwritten
  std::mt19937_64 rng(777);
  for (int it = 0; it < 600; ++it) {
    std::string b = good;
    int const k = 1 + static_cast<int>(rng() % 4);
    for (int j = 0; j < k; ++j) {
      std::size_t const pos = rng() % b.size();
      b[pos] = static_cast<char>(rng() & 0xff);
    }
    std::istringstream in(b);
    hgraph g;
    try {
      load(in, g);
      CHECK(hgraph_ok(g), "%s: corruption %d loaded inconsistent arrays", what, it);
    } catch (std::exception const&) {
    }
  }
}

void check_files() {
  for (uint32_t vpg : {1u, 4u}) {
    hgraph g = small_graph(vpg);
    CHECK(hgraph_ok(g), "generated graph");
    {
      std::ostringstream out;
      kites_amd::save_binary(g, out);
      std::string const bin = out.str();
      std::istringstream in(bin);
      hgraph h;
      kites_amd::load_binary(in, h);
      CHECK(h.R == g.R && h.C == g.C && h.E == g.E && h.oC == g.oC && h.dir_C == g.dir_C,
            "binary round trip");
      corrupt_and_load(bin, [](std::istream& i, hgraph& o) { kites_amd::load_binary(i, o); },
                       "binary", true);
    }
    if (vpg == 1) {
      std::ostringstream out;
      kites_amd::save_text(g, out);
      std::string const txt = out.str();
      std::istringstream in(txt);
      hgraph h;
      kites_amd::load_text(in, h, 2);
      CHECK(h.R == g.R && h.C == g.C && h.E == g.E, "text round trip");
      corrupt_and_load(txt, [](std::istream& i, hgraph& o) { kites_amd::load_text(i, o, 2); },
                       "text", false);
    }
  }
}

// ---- generators and CSR builds ---------------------------------------------

void check_generators() {
  for (uint32_t scale : {1u, 4u, 9u}) {
    hgraph g;
    kites_amd::gen_rmat(scale, 8, 5 + scale, 0.57, 0.19, 0.19, 1, 255, 0, 0, 3, g);
    CHECK(hgraph_ok(g) && g.n % 128 == 0, "rmat %u", scale);
    hgraph s;
    kites_amd::symmetrize(g, 2, s);
    CHECK(hgraph_ok(s), "symmetrize %u", scale);
  }
  {
    hgraph g;
    kites_amd::gen_rmat(13, 16, 1, 0.57, 0.19, 0.19, 1, 64, 7115, 103689, 4, g);
    CHECK(hgraph_ok(g) && g.m == 103689, "C1 graph m=%llu", static_cast<unsigned long long>(g.m));
  }
  {
    hgraph g;
    kites_amd::gen_er(1000, 5000, 3, 1, 10, 2, g);
    CHECK(hgraph_ok(g), "er");
  }
  for (uint32_t r : {1u, 2u, 17u}) {
    hgraph g;
    kites_amd::gen_grid(r, 9, 4, 1, 100, 2, g);
    CHECK(hgraph_ok(g), "grid %u x 9", r);
  }
  {
    hgraph g;
    std::vector<uint32_t> voff, src;
    kites_amd::gen_rmat_batch(12, 8, 3, 9, 2, 6, 1, 255, 3, g, voff, src);
    CHECK(hgraph_ok(g) && voff.size() == 13 && src.size() == 12, "batch");
    for (std::size_t i = 0; i < src.size(); ++i)
      CHECK(src[i] >= voff[i] && src[i] < voff[i + 1], "batch source %zu", i);
  }
  std::mt19937_64 rng(4);
  for (int it = 0; it < 50; ++it) {
    uint32_t const n = 1 + static_cast<uint32_t>(rng() % 300);
    std::size_t const m = rng() % 2000;
    std::vector<uint32_t> s(m), d(m), w(m);
    for (std::size_t e = 0; e < m; ++e) {
      s[e] = static_cast<uint32_t>(rng() % n);
      d[e] = static_cast<uint32_t>(rng() % n);
      w[e] = static_cast<uint32_t>(rng() % 50);
    }
    hgraph g;
    kites_amd::build_csr(kites_amd::pad_vertices(n), s, d, w, (it & 1) != 0, 2, g);
    CHECK(hgraph_ok(g), "build_csr %d", it);
  }
}

// ---- launch cache LRU -------------------------------------------------------

struct owned {
  int* p = nullptr;   // heap memory owned until released
};

void check_lru() {
  using kites::detail::byte_lru;
  int releases = 0;
  std::vector<int> released;
  auto rel = [&](owned& o) {
    CHECK(o.p != nullptr, "value released twice");
    released.push_back(*o.p);
    delete o.p;
    o.p = nullptr;
    ++releases;
  };
  auto key = [](int i, std::size_t len) {
    std::vector<unsigned char> k(len, 0);
    std::memcpy(k.data(), &i, std::min(len, sizeof(i)));
    return k;
  };
  {
    byte_lru<owned> c(3);
    for (int i = 0; i < 3; ++i) c.insert(key(i, 8), owned{new int(i)}, rel);
    CHECK(c.size() == 3 && releases == 0, "fill");
    CHECK(c.find(key(0, 8)) && *c.find(key(0, 8))->p == 0, "hit 0");
    CHECK(!c.find(key(0, 4)), "a key of another length is another key");
    c.insert(key(3, 8), owned{new int(3)}, rel);          // drops 1 (0 was used)
    CHECK(released.size() == 1 && released[0] == 1, "evicted %d", released.empty() ? -1 : released[0]);
    c.insert(key(4, 8), owned{new int(4)}, rel);          // drops 2
    c.insert(key(5, 8), owned{new int(5)}, rel);          // drops 0 (3 and 4 are newer)
    CHECK(released.size() == 3 && released[1] == 2 && released[2] == 0, "eviction order");
    // the value inserted last is never the next one dropped
    for (int i = 6; i < 40; ++i) {
      c.insert(key(i, 8), owned{new int(i)}, rel);
      CHECK(c.find(key(i, 8)) != nullptr, "fresh entry %d dropped", i);
      CHECK(c.size() <= 3, "capacity");
    }
    c.clear(rel);
    CHECK(c.size() == 0, "clear");
  }
  CHECK(releases == 40, "every value released once (%d)", releases);
  {
    // two handles taken back to back (a run's full- and small-grid chunks)
    // both stay cached whatever the fill state
    byte_lru<owned> c(8);
    std::mt19937_64 rng(9);
    for (int it = 0; it < 500; ++it) {
      int const a = static_cast<int>(rng() % 20), b = 100 + static_cast<int>(rng() % 20);
      for (int x : {a, b})
        if (!c.find(key(x, 16))) c.insert(key(x, 16), owned{new int(x)}, rel);
      CHECK(c.find(key(a, 16)) && c.find(key(b, 16)), "pair %d %d", a, b);
    }
    c.clear(rel);
  }
}

// ---- oracle -----------------------------------------------------------------

void check_oracle() {
  std::mt19937_64 rng(2024);
  for (int it = 0; it < 60; ++it) {
    uint32_t const n = 1 + static_cast<uint32_t>(rng() % 200);
    std::size_t const m = rng() % 1500;
    std::vector<uint32_t> s(m), d(m), w(m);
    for (std::size_t e = 0; e < m; ++e) {
      s[e] = static_cast<uint32_t>(rng() % n);
      d[e] = static_cast<uint32_t>(rng() % n);
      w[e] = 1 + static_cast<uint32_t>(rng() % 30);
    }
    hgraph g;
    kites_amd::build_csr(n, s, d, w, true, 1, g);
    uint32_t const src = static_cast<uint32_t>(rng() % n);
    std::vector<uint32_t> dj(n), V(n, 0xFFFFFFFFu), V2;
    oracle_dijkstra_u32(n, g.R.data(), g.C.data(), g.E.data(), src, dj.data());
    V[src] = 0;
    V2 = V;
    oracle_fixpoint_u32(n, g.R.data(), g.C.data(), g.E.data(), V.data(), KITES_ORACLE_SSSP_SAT, 0);
    CHECK(V == dj, "fixpoint vs dijkstra %d", it);
    std::vector<uint8_t> F(n, 0);
    for (uint32_t e = g.oR[src]; e < g.oR[src + 1]; ++e) F[g.oC[e]] = 1;
    int used = 0;
    oracle_frontier_omp_csr_u32(n, g.R.data(), g.C.data(), g.E.data(), g.oR.data(), g.oC.data(),
                                g.oE.data(), V2.data(), F.data(), KITES_ORACLE_SSSP_SAT, 2, &used);
    CHECK(V2 == dj, "frontier vs dijkstra %d", it);
    uint64_t bad[3];
    uint32_t const srcs[1] = {src};
    CHECK(oracle_check_cols_u32(n, g.R.data(), g.C.data(), g.E.data(), 1, dj.data(), srcs, 2,
                                bad) == 0,
          "certificate %d", it);
    std::vector<uint32_t> lv(n), lab(n);
    oracle_bfs_u32(n, g.R.data(), g.C.data(), src, lv.data());
    oracle_cc_min_label(n, g.R.data(), g.C.data(), lab.data());
    for (uint32_t v = 0; v < n; ++v) CHECK(lab[v] <= v, "cc label above the id");
  }
}

}  // namespace

int main() {
  check_lru();
  check_parser();
  check_files();
  check_generators();
  check_oracle();
  if (g_fail) {
    std::fprintf(stderr, "%d host checks failed\n", g_fail);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
