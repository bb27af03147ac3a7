// graph.hpp -- host graph construction: edge-list parser and the pinned
// host CSR graph.
//
// graph_raw restates the reference parser's behaviour
// (kites/graph/graph_raw.h:58-190): lines whose first character is not a
// digit are skipped; tokens are separated by space or tab only (so the '\r'
// of a CRLF line stays on the last token and the integer conversion stops
// there); tokens after the second are passed to the pre-processing callback
// as argv; every edge becomes an in-neighbour of its destination; the
// callback's return value sets the initial frontier bit of the destination's
// group; in-lists are sorted by source id with duplicates kept; directed
// graphs get group-level out-lists, de-duplicated when vpg != 1; the vertex
// count is padded to a multiple of 128.
//
// graph_csr keeps the reference's public members and constructors
// (kites/graph/graph_csr.h:21-232) and adds one constructor from in-memory
// CSR arrays.  One deliberate difference: for undirected graphs the
// reference never copies the callback's frontier bits into vBitmap (its copy
// loop runs over the directed-only out-lists, graph_csr.h:100-108), so its
// GPU run starts from an empty frontier and fails its own sanity_check; here
// the bits are kept for undirected graphs too.
#ifndef KITES_GRAPH_HPP_
#define KITES_GRAPH_HPP_

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"
#include "io.hpp"
#include "memory.hpp"

namespace kites {

class graph_base {
 protected:
  std::size_t numVertices;
  std::size_t numEdges;

 public:
  graph_base(std::size_t nV = 0, std::size_t nE = 0) : numVertices(nV), numEdges(nE) {}
  void set_num_vertices(std::size_t n) { numVertices = n; }
  void set_num_edges(std::size_t m) { numEdges = m; }
  std::size_t get_num_vertices() const { return numVertices; }
  std::size_t get_num_edges() const { return numEdges; }
};

namespace detail {

// C atoll: leading white space, optional sign, decimal digits.
inline long long parse_ll(const char* s) {
  while (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\v' || *s == '\f' || *s == '\r')
    ++s;
  bool neg = false;
  if (*s == '+' || *s == '-') neg = (*s++ == '-');
  unsigned long long v = 0;
  while (*s >= '0' && *s <= '9') v = v * 10 + (unsigned)(*s++ - '0');
  return neg ? -(long long)v : (long long)v;
}

// Splits `line` in place at spaces and tabs (each separator becomes a NUL;
// the string's own terminator ends the last token); returns token pointers.
inline void split_tokens(std::string& line, std::vector<char*>& toks) {
  toks.clear();
  char* p = &line[0];
  char* end = p + line.size();
  while (p < end) {
    while (p < end && (*p == ' ' || *p == '\t')) *p++ = '\0';
    if (p >= end) break;
    toks.push_back(p);
    while (p < end && *p != ' ' && *p != '\t') ++p;
  }
}

// Checks one CSR of a prepared graph read from an untrusted file before
// anything indexes with it: offsets from 0 to m, never decreasing, every id
// below n.  Throws std::runtime_error naming `what`.
template <class I, class J>
void check_prepared_csr(std::size_t n, std::size_t m, const I* R, const J* C, const char* what) {
  if (static_cast<std::size_t>(R[0]) != 0 || static_cast<std::size_t>(R[n]) != m)
    throw std::runtime_error(std::string("Prepared graph: ") + what +
                             " offsets do not span the edges.");
  for (std::size_t v = 0; v < n; ++v)
    if (R[v] > R[v + 1])
      throw std::runtime_error(std::string("Prepared graph: ") + what + " offsets decrease.");
  for (std::size_t e = 0; e < m; ++e)
    if (static_cast<std::size_t>(C[e]) >= n)
      throw std::runtime_error(std::string("Prepared graph: ") + what + " id out of range.");
}

}  // namespace detail

template <typename vT, class eT, class vIdxT = unsigned int>
class graph_raw : public graph_base {
 public:
  struct neighbor {
    eT EdgeVal{};
    vIdxT nbrIdx;
    explicit neighbor(vIdxT i) : nbrIdx{i} {}
    eT& getEdgeVal() { return EdgeVal; }
    bool operator<(neighbor const& r) const { return nbrIdx < r.nbrIdx; }
    bool operator==(neighbor const& r) const { return nbrIdx == r.nbrIdx; }
  };

  std::vector<vT> rawVVec;
  std::vector<std::vector<neighbor>> rawNbrsVec;
  std::vector<std::vector<vIdxT>> rawIVecOutgoing;
  std::vector<bool> vBitmap;
  unsigned vertexPerGroup;
  std::size_t totalNumOutEdges;

 private:
  template <class funcT>
  void construct_me(std::istream& in, funcT f, input_graph_form gForm,
                    graph_property gProp, edge_list_expression elExp) {
    std::string line;
    std::vector<char*> toks;
    const unsigned vpg = vertexPerGroup;
    while (std::getline(in, line)) {
      if (line.empty() || line[0] < '0' || line[0] > '9') continue;
      detail::split_tokens(line, toks);
      if (toks.size() < 2) continue;
      vIdxT const a = static_cast<vIdxT>(detail::parse_ll(toks[0]));
      vIdxT const b = static_cast<vIdxT>(detail::parse_ll(toks[1]));
      vIdxT const mx = a >= b ? a : b;
      vIdxT const s = (gForm == input_graph_form::edge_list_s_d) ? a : b;
      vIdxT const d = (gForm == input_graph_form::edge_list_s_d) ? b : a;
      if (rawVVec.size() <= static_cast<std::size_t>(mx)) {
        rawNbrsVec.resize(static_cast<std::size_t>(mx) + 1);
        rawVVec.resize(static_cast<std::size_t>(mx) + 1);
        vBitmap.resize((static_cast<std::size_t>(mx) + 1 + vpg - 1) / vpg, false);
      }
      unsigned argc = static_cast<unsigned>(std::min<std::size_t>(toks.size() - 2, 60));
      char** argv = toks.data() + 2;
      rawNbrsVec[d].push_back(neighbor(s));
      bool r = f(argc, argv, s, d, rawVVec[s], rawVVec[d], rawNbrsVec[d].back().getEdgeVal());
      if (r) vBitmap[d / vpg] = true;
      if (gProp == graph_property::undirected && elExp == edge_list_expression::incomplete) {
        rawNbrsVec[s].push_back(neighbor(d));
        bool r2 = f(argc, argv, d, s, rawVVec[d], rawVVec[s],
                    rawNbrsVec[s].back().getEdgeVal());
        if (r2) vBitmap[s / vpg] = true;
      }
    }

    set_num_edges(0);
    std::size_t m = 0;
    if (gProp == graph_property::directed)
      rawIVecOutgoing.resize((rawVVec.size() + vpg - 1) / vpg);
    for (std::size_t dst = 0; dst < rawNbrsVec.size(); ++dst) {
      auto& lst = rawNbrsVec[dst];
      std::stable_sort(lst.begin(), lst.end());
      m += lst.size();
      if (gProp == graph_property::directed)
        for (auto& nb : lst) rawIVecOutgoing[nb.nbrIdx / vpg].push_back(static_cast<vIdxT>(dst / vpg));
    }
    set_num_edges(m);
    if (gProp == graph_property::directed && vpg != 1) {
      totalNumOutEdges = 0;
      for (auto& el : rawIVecOutgoing) {
        std::sort(el.begin(), el.end());
        el.erase(std::unique(el.begin(), el.end()), el.end());
        totalNumOutEdges += el.size();
      }
    } else {
      totalNumOutEdges = m;
    }
    std::size_t const n0 = rawVVec.size();
    if (n0 % VERTEX_PADDING != 0) {
      std::size_t const np = (n0 / VERTEX_PADDING + 1) * VERTEX_PADDING;
      rawNbrsVec.resize(np);
      rawVVec.resize(np);
      if (gProp == graph_property::directed && vpg == 1) {
        rawIVecOutgoing.resize(np);
        vBitmap.resize(np);
      }
    }
    set_num_vertices(rawVVec.size());
  }

 public:
  template <class funcT>
  graph_raw(std::istream& in, funcT f, unsigned vpg,
            input_graph_form gForm = input_graph_form::edge_list_s_d,
            graph_property gProp = graph_property::directed,
            edge_list_expression elExp = edge_list_expression::complete)
      : graph_base{}, vertexPerGroup{vpg == 0 ? 1u : vpg}, totalNumOutEdges{0} {
    construct_me(in, f, gForm, gProp, elExp);
  }
  template <class funcT>
  graph_raw(std::string const& file, funcT f, unsigned vpg,
            input_graph_form gForm = input_graph_form::edge_list_s_d,
            graph_property gProp = graph_property::directed,
            edge_list_expression elExp = edge_list_expression::complete)
      : graph_base{}, vertexPerGroup{vpg == 0 ? 1u : vpg}, totalNumOutEdges{0} {
    std::ifstream in;
    io::openFileToAccess(in, file);
    construct_me(in, f, gForm, gProp, elExp);
  }
  unsigned getVertexPerGroup() const { return vertexPerGroup; }
  std::size_t getTotalNumOutEdges() const { return totalNumOutEdges; }
};

// Pinned host graph in CSR-of-in-edges form (graph_csr.h:21-232).
template <class vT, class eT, class vIdxT = unsigned int>
class graph_csr : public graph_base {
  using eIdxT = vIdxT;

 public:
  host_pinned_buffer<vT> V;
  host_pinned_buffer<eIdxT> R;
  host_pinned_buffer<vIdxT> C;
  host_pinned_buffer<eT> E;
  host_bitmap vBitmap;
  host_pinned_buffer<unsigned> updateFlag;
  graph_property gProp;
  unsigned vertexPerGroup;
  host_pinned_buffer<eIdxT> dir_R;
  host_pinned_buffer<vIdxT> dir_C;

 private:
  template <class funcT>
  void construct_me(std::istream& in, funcT f, unsigned vpg, input_graph_form gForm,
                    graph_property gp, edge_list_expression elExp) {
    graph_raw<vT, eT, vIdxT> g(in, f, vpg, gForm, gp, elExp);
    std::size_t const n = g.get_num_vertices(), m = g.get_num_edges();
    set_num_vertices(n);
    set_num_edges(m);
    V.alloc(n);
    R.alloc(n + 1);
    C.alloc(m);
    E.alloc(m);
    updateFlag.alloc(1);
    updateFlag[0] = 0;
    vBitmap.alloc(g.vBitmap.size());
    vBitmap.reset();
    if (n) std::memcpy(V.get_ptr(), g.rawVVec.data(), n * sizeof(vT));
    R[0] = 0;
    for (std::size_t v = 0; v < n; ++v) {
      auto const& lst = g.rawNbrsVec[v];
      eIdxT off = R[v];
      for (std::size_t k = 0; k < lst.size(); ++k) {
        C[off + k] = lst[k].nbrIdx;
        E[off + k] = lst[k].EdgeVal;
      }
      R[v + 1] = static_cast<eIdxT>(off + lst.size());
    }
    for (std::size_t b = 0; b < g.vBitmap.size(); ++b)
      if (g.vBitmap[b]) vBitmap.setAt(b);
    if (gp == graph_property::directed) {
      dir_R.alloc(g.rawIVecOutgoing.size() + 1);
      dir_C.alloc(g.getTotalNumOutEdges());
      dir_R[0] = 0;
      std::size_t k = 0;
      for (std::size_t gi = 0; gi < g.rawIVecOutgoing.size(); ++gi) {
        for (auto x : g.rawIVecOutgoing[gi]) dir_C[k++] = x;
        dir_R[gi + 1] = static_cast<eIdxT>(k);
      }
    }
  }

  // Reads the text dump written by io::save_to_file (graph_csr.h:112-166).
  void construct_prepared(std::istream& in) {
    std::string line;
    auto next = [&]() -> unsigned long long {
      if (!std::getline(in, line)) throw std::runtime_error("Truncated prepared graph file.");
      return std::stoull(line);
    };
    std::size_t const n = next(), m = next();
    gProp = (next() == 1) ? graph_property::directed : graph_property::undirected;
    // The file is untrusted: sizes must fit the index types, and the
    // offsets and ids are checked below before the device build indexes
    // with them.
    if (n >= 0xffffffffull || m > 0xffffffffull)
      throw std::runtime_error("Prepared graph exceeds 32-bit ids or offsets.");
    set_num_vertices(n);
    set_num_edges(m);
    V.alloc(n);
    R.alloc(n + 1);
    C.alloc(m);
    E.alloc(m);
    vBitmap.alloc(n);
    updateFlag.alloc(1);
    updateFlag[0] = 0;
    for (std::size_t i = 0; i < n; ++i) V[i] = static_cast<vT>(next());
    for (std::size_t i = 0; i < n + 1; ++i) R[i] = static_cast<eIdxT>(next());
    for (std::size_t i = 0; i < m; ++i) C[i] = static_cast<vIdxT>(next());
    for (std::size_t i = 0; i < m; ++i) E[i] = static_cast<eT>(next());
    for (std::size_t i = 0; i < vBitmap.size(); ++i) vBitmap[i] = static_cast<unsigned>(next());
    if (gProp == graph_property::directed) {
      dir_R.alloc(n + 1);
      dir_C.alloc(m);
      for (std::size_t i = 0; i < n + 1; ++i) dir_R[i] = static_cast<eIdxT>(next());
      for (std::size_t i = 0; i < m; ++i) dir_C[i] = static_cast<vIdxT>(next());
    }
    detail::check_prepared_csr(n, m, R.get_ptr(), C.get_ptr(), "in-edge");
    if (gProp == graph_property::directed)
      detail::check_prepared_csr(n, m, dir_R.get_ptr(), dir_C.get_ptr(), "out-list");
  }

 public:
  unsigned getVertexPerGroup() const { return vertexPerGroup; }

  template <class funcT>
  graph_csr(std::ifstream& in, funcT f, input_graph_form gForm = input_graph_form::edge_list_s_d,
            graph_property gp = graph_property::directed,
            edge_list_expression elExp = edge_list_expression::complete)
      : gProp{gp}, vertexPerGroup{1} {
    construct_me(in, f, 1, gForm, gp, elExp);
  }
  template <class funcT>
  graph_csr(std::string const& file, funcT f,
            input_graph_form gForm = input_graph_form::edge_list_s_d,
            graph_property gp = graph_property::directed,
            edge_list_expression elExp = edge_list_expression::complete)
      : gProp{gp}, vertexPerGroup{1} {
    std::ifstream in;
    io::openFileToAccess(in, file);
    construct_me(in, f, 1, gForm, gp, elExp);
  }
  explicit graph_csr(std::ifstream& in) : vertexPerGroup{1} { construct_prepared(in); }
  explicit graph_csr(std::string const& file) : vertexPerGroup{1} {
    std::ifstream in;
    io::openFileToAccess(in, file);
    construct_prepared(in);
  }
  template <class funcT>
  graph_csr(std::ifstream& in, funcT f, unsigned vpg,
            input_graph_form gForm = input_graph_form::edge_list_s_d)
      : gProp{graph_property::directed}, vertexPerGroup{vpg} {
    construct_me(in, f, vpg, gForm, graph_property::directed, edge_list_expression::complete);
  }
  template <class funcT>
  graph_csr(std::string const& file, funcT f, unsigned vpg,
            input_graph_form gForm = input_graph_form::edge_list_s_d)
      : gProp{graph_property::directed}, vertexPerGroup{vpg} {
    std::ifstream in;
    io::openFileToAccess(in, file);
    construct_me(in, f, vpg, gForm, graph_property::directed, edge_list_expression::complete);
  }

  // Extension: from in-memory in-edge CSR arrays (R: n+1, C/E: m), initial
  // values V0 (n) and initial frontier bits (n bytes, nonzero = active).
  // Vertex-per-group is 1; directed graphs also get dir_R / dir_C.
  graph_csr(std::size_t n, const eIdxT* Rin, const vIdxT* Cin, const eT* Ein,
            const vT* V0, const unsigned char* frontier0,
            graph_property gp = graph_property::directed)
      : gProp{gp}, vertexPerGroup{1} {
    std::size_t const m = n ? Rin[n] : 0;
    set_num_vertices(n);
    set_num_edges(m);
    V.alloc(n);
    R.alloc(n + 1);
    C.alloc(m);
    E.alloc(m);
    updateFlag.alloc(1);
    updateFlag[0] = 0;
    vBitmap.alloc(n);
    vBitmap.reset();
    if (n) {
      std::memcpy(V.get_ptr(), V0, n * sizeof(vT));
      std::memcpy(R.get_ptr(), Rin, (n + 1) * sizeof(eIdxT));
    }
    if (m) {
      std::memcpy(C.get_ptr(), Cin, m * sizeof(vIdxT));
      std::memcpy(E.get_ptr(), Ein, m * sizeof(eT));
    }
    for (std::size_t v = 0; v < n; ++v)
      if (frontier0 && frontier0[v]) vBitmap.setAt(v);
    if (gp == graph_property::directed) {
      dir_R.alloc(n + 1);
      dir_C.alloc(m);
      std::vector<std::size_t> cnt(n + 1, 0);
      for (std::size_t e = 0; e < m; ++e) cnt[C[e] + 1]++;
      for (std::size_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
      for (std::size_t v = 0; v <= n; ++v) dir_R[v] = static_cast<eIdxT>(cnt[v]);
      for (std::size_t v = 0; v < n; ++v)
        for (eIdxT e = R[v]; e < R[v + 1]; ++e) dir_C[cnt[C[e]]++] = static_cast<vIdxT>(v);
    }
  }
};

}  // namespace kites

#endif  // KITES_GRAPH_HPP_
