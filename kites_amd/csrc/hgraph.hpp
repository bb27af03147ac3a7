// hgraph.hpp -- host CSR graph used by the C-ABI (generators, edge-list
// loading, upload).  Plain C++; no HIP.
#ifndef KITES_AMD_HGRAPH_HPP_
#define KITES_AMD_HGRAPH_HPP_

#include <cstdint>
#include <iosfwd>
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

// `count` independent R-MAT graphs (a=.57, b=.19, c=.19) with scale and edge
// factor drawn uniformly from [scale_lo, scale_hi] x [ef_lo, ef_hi], laid
// out as one disjoint union: graph i owns vertices [voff[i], voff[i+1]) and
// a contiguous range of the in- and out-edge arrays.  sources[i] = graph
// i's vertex of maximum out-degree (lowest id on ties), as a union id.
void gen_rmat_batch(uint32_t count, uint64_t seed, uint32_t scale_lo, uint32_t scale_hi,
                    uint32_t ef_lo, uint32_t ef_hi, uint32_t wmin, uint32_t wmax, int threads,
                    hgraph& g, std::vector<uint32_t>& voff, std::vector<uint32_t>& sources);

// Prepared-graph files (graph_files.cpp): the reference's text dump and
// reader (kites/io/graph_output.cuh:29-35, kites/graph/graph_csr.h:112-166)
// and a raw binary format that also holds the out-edge CSR.  Readers throw
// std::runtime_error on truncated or inconsistent input.
void save_text(const hgraph& g, std::ostream& out);
void load_text(std::istream& in, hgraph& g, int threads);
void save_binary(const hgraph& g, std::ostream& out);
void load_binary(std::istream& in, hgraph& g);

}  // namespace kites_amd

#endif  // KITES_AMD_HGRAPH_HPP_
