// graph_files.cpp -- prepared-graph files for host graphs.
//
// Text: the reference's dump (kites/io/graph_output.cuh:29-35, each buffer
// written one value per line as host_pinned_buffer's operator<< does,
// kites/memory/host_pinned_buffer.cuh:120-124) and its reader
// (kites/graph/graph_csr.h:112-166):
//   n, m, directed (1/0); V (n); R (n+1); C (m); E (m); the frontier bitmap
//   as 32-bit words (bit v of word v/32, host_pinned_buffer.cuh:151-153);
//   directed only: dir_R, dir_C (the out-lists).
// The reader, like the reference's, assumes one vertex per group: it reads
// ceil(n/32) bitmap words, n+1 dir_R and m dir_C entries.
//
// Binary (no counterpart in the reference): the same arrays plus the
// weighted out-edge CSR, raw little-endian, so loading is a few reads with
// no parsing and no transpose.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "hgraph.hpp"

namespace kites_amd {

namespace {

template <class T>
void put_lines(std::ostream& out, const T* a, std::size_t n) {
  std::string buf;
  buf.reserve(1 << 20);
  char tmp[24];
  for (std::size_t i = 0; i < n; ++i) {
    int const k = std::snprintf(tmp, sizeof(tmp), "%llu\n", static_cast<unsigned long long>(a[i]));
    buf.append(tmp, static_cast<std::size_t>(k));
    if (buf.size() > (1u << 20) - 32) {
      out.write(buf.data(), static_cast<std::streamsize>(buf.size()));
      buf.clear();
    }
  }
  out.write(buf.data(), static_cast<std::streamsize>(buf.size()));
}

// One line as std::stoull reads it (leading white space, decimal digits);
// a missing or empty line is a truncated file.
struct line_reader {
  std::istream& in;
  std::string line;
  unsigned long long next() {
    if (!std::getline(in, line)) throw std::runtime_error("Truncated prepared graph file.");
    std::size_t pos = 0;
    unsigned long long v = 0;
    try {
      v = std::stoull(line, &pos);
    } catch (const std::exception&) {
      throw std::runtime_error("Malformed prepared graph file: '" + line + "'");
    }
    return v;
  }
  template <class T>
  void fill(std::vector<T>& a, std::size_t n) {
    a.resize(n);
    for (std::size_t i = 0; i < n; ++i) a[i] = static_cast<T>(next());
  }
};

constexpr char kMagic[8] = {'K', 'I', 'T', 'E', 'S', 'G', 'R', '1'};

struct bin_header {
  char magic[8];
  uint32_t n, n_real, vpg, directed;
  uint64_t m, nV0, nfront, ndirR, ndirC;
};

template <class T>
void put_raw(std::ostream& out, const std::vector<T>& a) {
  if (!a.empty())
    out.write(reinterpret_cast<const char*>(a.data()), static_cast<std::streamsize>(a.size() * sizeof(T)));
}

template <class T>
void get_raw(std::istream& in, std::vector<T>& a, uint64_t n) {
  a.resize(n);
  if (n && !in.read(reinterpret_cast<char*>(a.data()), static_cast<std::streamsize>(n * sizeof(T))))
    throw std::runtime_error("Truncated binary graph file.");
}

// Offsets from 0 to m, never decreasing, ids below n (untrusted input).
void check_csr(uint64_t n, uint64_t m, const std::vector<uint32_t>& R,
               const std::vector<uint32_t>& C, const char* span, const char* dec,
               const char* range) {
  auto fail = [](const char* w) { throw std::runtime_error(std::string("Prepared graph: ") + w + "."); };
  if (R[0] != 0 || R[n] != m) fail(span);
  for (uint64_t v = 0; v < n; ++v)
    if (R[v] > R[v + 1]) fail(dec);
  for (uint64_t e = 0; e < m; ++e)
    if (C[e] >= n) fail(range);
}

}  // namespace

void save_text(const hgraph& g, std::ostream& out) {
  uint64_t const n = g.n, m = g.m;
  out << n << "\n" << m << "\n" << (g.directed ? 1 : 0) << "\n";
  if (g.V0.size() == n) {
    put_lines(out, g.V0.data(), n);
  } else {
    std::vector<uint32_t> z(n, 0);
    put_lines(out, z.data(), n);
  }
  put_lines(out, g.R.data(), n + 1);
  put_lines(out, g.C.data(), m);
  put_lines(out, g.E.data(), m);
  // Frontier bits per group; the reference sets them only in its loop over
  // the out-list groups, which is empty for undirected graphs
  // (graph_csr.h:96-104).
  uint64_t const nbits = g.frontier0.empty() ? n : g.frontier0.size();
  uint64_t const ngroups_out =
      g.directed ? (g.dir_R.size() > 1 ? g.dir_R.size() - 1 : n) : 0;
  std::vector<uint32_t> words((nbits + 31) / 32, 0);
  for (uint64_t i = 0; i < g.frontier0.size() && i < ngroups_out; ++i)
    if (g.frontier0[i]) words[i >> 5] |= 1u << (i & 31);
  put_lines(out, words.data(), words.size());
  if (g.directed) {
    if (g.dir_R.size() > 1) {
      put_lines(out, g.dir_R.data(), g.dir_R.size());
      put_lines(out, g.dir_C.data(), g.dir_C.size());
    } else {
      // vertex-per-group 1: the out-lists are the out-edge CSR's
      put_lines(out, g.oR.data(), n + 1);
      put_lines(out, g.oC.data(), m);
    }
  }
  if (!out) throw std::runtime_error("Writing the prepared graph file failed.");
}

void load_text(std::istream& in, hgraph& g, int threads) {
  line_reader r{in, {}};
  uint64_t const n = r.next();
  uint64_t const m = r.next();
  bool const directed = r.next() == 1;
  if (n >= 0xffffffffull || m >= 0xffffffffull)
    throw std::runtime_error("Prepared graph exceeds 32-bit ids or offsets.");
  g = hgraph{};
  g.n = static_cast<uint32_t>(n);
  g.n_real = g.n;
  g.m = m;
  g.directed = directed;
  g.vpg = 1;
  r.fill(g.V0, n);
  r.fill(g.R, n + 1);
  r.fill(g.C, m);
  r.fill(g.E, m);
  std::vector<uint32_t> words;
  r.fill(words, (n + 31) / 32);
  g.frontier0.assign(n, 0);
  for (uint64_t v = 0; v < n; ++v) g.frontier0[v] = (words[v >> 5] >> (v & 31)) & 1u;
  if (directed) {
    r.fill(g.dir_R, n + 1);
    r.fill(g.dir_C, m);
  } else {
    g.dir_R.assign(1, 0);
  }
  check_csr(n, m, g.R, g.C, "R does not span m edges", "R decreases", "source id out of range");
  if (directed)
    check_csr(n, m, g.dir_R, g.dir_C, "dir_R does not span m edges", "dir_R decreases",
              "out-list id out of range");
  transpose(g, threads);
}

void save_binary(const hgraph& g, std::ostream& out) {
  bin_header h{};
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.n = g.n;
  h.n_real = g.n_real;
  h.vpg = g.vpg;
  h.directed = g.directed ? 1u : 0u;
  h.m = g.m;
  h.nV0 = g.V0.size();
  h.nfront = g.frontier0.size();
  h.ndirR = g.dir_R.size();
  h.ndirC = g.dir_C.size();
  out.write(reinterpret_cast<const char*>(&h), sizeof(h));
  put_raw(out, g.R);
  put_raw(out, g.C);
  put_raw(out, g.E);
  put_raw(out, g.oR);
  put_raw(out, g.oC);
  put_raw(out, g.oE);
  put_raw(out, g.V0);
  put_raw(out, g.frontier0);
  put_raw(out, g.dir_R);
  put_raw(out, g.dir_C);
  if (!out) throw std::runtime_error("Writing the binary graph file failed.");
}

void load_binary(std::istream& in, hgraph& g) {
  bin_header h{};
  if (!in.read(reinterpret_cast<char*>(&h), sizeof(h)) ||
      std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0)
    throw std::runtime_error("Not a binary graph file.");
  // The header is untrusted: sizes are checked against the 32-bit id and
  // offset types and against the bytes the file holds before anything is
  // allocated, and the vertex-group lists (uploaded for vertex-per-group
  // > 1 and walked by the iteration kernel) are checked entry by entry.
  if (h.n >= 0xffffffffull || h.m > 0xffffffffull || h.n_real > h.n)
    throw std::runtime_error("Binary graph file: sizes exceed 32-bit ids or offsets.");
  if (h.vpg == 0 || (h.vpg & (h.vpg - 1)) != 0)
    throw std::runtime_error("Binary graph file: vertex-per-group is not a power of two.");
  uint64_t const ngroups = (static_cast<uint64_t>(h.n) + h.vpg - 1) / h.vpg;
  if (h.nV0 != 0 && h.nV0 != h.n)
    throw std::runtime_error("Binary graph file: initial values do not cover the vertices.");
  if (h.nfront > h.n) throw std::runtime_error("Binary graph file: frontier longer than n.");
  if (h.vpg > 1 && (h.ndirR == 0 || h.ndirR > ngroups + 1))
    throw std::runtime_error("Binary graph file: group out-offsets do not match the groups.");
  if (h.ndirR > static_cast<uint64_t>(h.n) + 1 || h.ndirC > 0xffffffffull)
    throw std::runtime_error("Binary graph file: group lists too long.");
  {
    uint64_t const need = 4 * (2 * (static_cast<uint64_t>(h.n) + 1) + 4 * h.m + h.nV0 +
                               h.ndirR + h.ndirC) + h.nfront;
    std::streampos const here = in.tellg();
    if (here != std::streampos(-1) && in.seekg(0, std::ios::end)) {
      std::streampos const end = in.tellg();
      in.seekg(here);
      if (end != std::streampos(-1) && static_cast<uint64_t>(end - here) < need)
        throw std::runtime_error("Truncated binary graph file.");
    }
    in.clear();
    in.seekg(here);
  }
  g = hgraph{};
  g.n = h.n;
  g.n_real = h.n_real;
  g.vpg = h.vpg;
  g.directed = h.directed != 0;
  g.m = h.m;
  get_raw(in, g.R, static_cast<uint64_t>(h.n) + 1);
  get_raw(in, g.C, h.m);
  get_raw(in, g.E, h.m);
  get_raw(in, g.oR, static_cast<uint64_t>(h.n) + 1);
  get_raw(in, g.oC, h.m);
  get_raw(in, g.oE, h.m);
  get_raw(in, g.V0, h.nV0);
  get_raw(in, g.frontier0, h.nfront);
  get_raw(in, g.dir_R, h.ndirR);
  get_raw(in, g.dir_C, h.ndirC);
  if (g.R[0] != 0 || g.R[h.n] != h.m || g.oR[0] != 0 || g.oR[h.n] != h.m)
    throw std::runtime_error("Binary graph file: offsets do not span m edges.");
  for (uint64_t v = 0; v < h.n; ++v)
    if (g.R[v] > g.R[v + 1] || g.oR[v] > g.oR[v + 1])
      throw std::runtime_error("Binary graph file: offsets decrease.");
  for (uint64_t e = 0; e < h.m; ++e)
    if (g.C[e] >= h.n || g.oC[e] >= h.n)
      throw std::runtime_error("Binary graph file: vertex id out of range.");
  if (!g.dir_R.empty()) {
    if (g.dir_R[0] != 0 || g.dir_R.back() != g.dir_C.size())
      throw std::runtime_error("Binary graph file: group out-offsets do not span the group lists.");
    for (std::size_t i = 0; i + 1 < g.dir_R.size(); ++i)
      if (g.dir_R[i] > g.dir_R[i + 1])
        throw std::runtime_error("Binary graph file: group out-offsets decrease.");
    uint64_t const bound = h.vpg > 1 ? ngroups : h.n;
    for (uint32_t x : g.dir_C)
      if (x >= bound) throw std::runtime_error("Binary graph file: group id out of range.");
  } else if (!g.dir_C.empty()) {
    throw std::runtime_error("Binary graph file: group lists without offsets.");
  }
}

}  // namespace kites_amd
