// hgraph.hpp -- host CSR graph used by the C-ABI (generators, edge-list
// loading, upload).  Plain C++; no HIP.
#ifndef KITES_AMD_HGRAPH_HPP_
#define KITES_AMD_HGRAPH_HPP_

#include <cstdint>
#include <vector>

namespace kites_amd {

struct hgraph {
  uint32_t n = 0;              // vertices (padded)
  uint32_t n_real = 0;         // vertices before padding
  uint64_t m = 0;
  std::vector<uint32_t> R, C, E;     // in-edge CSR: in-list of v = C[R[v]..R[v+1])
  std::vector<uint32_t> oR, oC, oE;  // out-edge CSR (transpose), weighted
  // Edge-list loads only: initial values, initial frontier bits (per group)
  // and group out-lists as the reference builds them.
  std::vector<uint32_t> V0;
  std::vector<uint8_t> frontier0;
  std::vector<uint32_t> dir_R, dir_C;
  uint32_t vpg = 1;
  bool directed = true;
};

// splitmix64 (Steele, Lea, Flood: "Fast splittable pseudorandom number
// generators", OOPSLA 2014): the counter-based generator behind every
// synthetic graph, so each graph is a pure function of its parameters.
inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Builds in- and out-CSR from edge arrays (src, dst, w), keeping duplicates
// and self-loops out when `dedup` is set (the first-drawn copy of a
// duplicate (src, dst) pair keeps its weight).  `n` must cover all ids.
void build_csr(uint32_t n, std::vector<uint32_t>& src, std::vector<uint32_t>& dst,
               std::vector<uint32_t>& w, bool dedup, int threads, hgraph& g);

void transpose(hgraph& g, int threads);

void gen_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed, double a, double b, double c,
              uint32_t wmin, uint32_t wmax, uint32_t n_limit, uint64_t exact_edges, int threads,
              hgraph& g);
void gen_er(uint32_t n, uint64_t draws, uint64_t seed, uint32_t wmin, uint32_t wmax, int threads,
            hgraph& g);
void gen_grid(uint32_t rows, uint32_t cols, uint64_t seed, uint32_t wmin, uint32_t wmax,
              int threads, hgraph& g);
void symmetrize(const hgraph& in, int threads, hgraph& out);

uint32_t pad_vertices(uint32_t n);

}  // namespace kites_amd

#endif  // KITES_AMD_HGRAPH_HPP_
