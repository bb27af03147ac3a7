// graph_raw_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Links the reference's own host-side edge-list parser
// (kites/graph/graph_raw.h, included from the read-only reference checkout at
// build time; no reference source is copied here) into a tiny CLI that dumps
// what graph_raw builds, so tests can pin the oracle's restatement of the
// parser and the product's C++ parser against the reference itself.
//
// usage: graph_raw_driver <edgelist> <out.bin> <vpg> <form:1=s_d,0=d_s>
//                         <prop:1=directed,0=undirected>
//                         <expr:1=complete,0=incomplete> <source>
// Callback: the reference example's pre-processing function semantics
// (examples/sssp_basic_single_device.cu:10-26) with a configurable source.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <vector>
#include "kites/graph/graph_raw.h"

static unsigned g_source = 30;

static bool example_callback(const unsigned argc, char** argv, const unsigned s,
                             const unsigned d, unsigned& sv, unsigned& dv,
                             unsigned& w) {
  sv = (s != g_source) ? 1073741823u : 0u;
  dv = (d != g_source) ? 1073741823u : 0u;
  w = (argc > 0) ? (unsigned)atoi(argv[0]) : ((s + 1) * (d + 2) % 127);
  return s == g_source;
}

static void put_u64(FILE* f, uint64_t x) { fwrite(&x, 8, 1, f); }
static void put_u32s(FILE* f, const std::vector<uint32_t>& v) {
  put_u64(f, v.size());
  if (!v.empty()) fwrite(v.data(), 4, v.size(), f);
}

int main(int argc, char** argv) {
  if (argc != 8) { std::fprintf(stderr, "bad args\n"); return 2; }
  unsigned vpg = (unsigned)atoi(argv[3]);
  auto form = atoi(argv[4]) ? kites::input_graph_form::edge_list_s_d
                            : kites::input_graph_form::edge_list_d_s;
  auto prop = atoi(argv[5]) ? kites::graph_property::directed
                            : kites::graph_property::undirected;
  auto expr = atoi(argv[6]) ? kites::edge_list_expression::complete
                            : kites::edge_list_expression::incomplete;
  g_source = (unsigned)atoi(argv[7]);
  kites::graph_raw<unsigned, unsigned> g(std::string(argv[1]), example_callback,
                                         vpg, form, prop, expr);
  std::vector<uint32_t> V(g.rawVVec.begin(), g.rawVVec.end());
  std::vector<uint32_t> R{0}, C, E;
  for (auto& lst : g.rawNbrsVec) {
    for (auto& nb : lst) { C.push_back(nb.nbrIdx); E.push_back(nb.EdgeVal); }
    R.push_back((uint32_t)C.size());
  }
  std::vector<uint32_t> B;
  for (bool b : g.vBitmap) B.push_back(b ? 1u : 0u);
  std::vector<uint32_t> dR{0}, dC;
  for (auto& lst : g.rawIVecOutgoing) {
    for (auto x : lst) dC.push_back(x);
    dR.push_back((uint32_t)dC.size());
  }
  FILE* f = std::fopen(argv[2], "wb");
  if (!f) return 3;
  put_u64(f, g.get_num_vertices());
  put_u64(f, g.get_num_edges());
  put_u64(f, g.getTotalNumOutEdges());
  put_u32s(f, V); put_u32s(f, R); put_u32s(f, C); put_u32s(f, E);
  put_u32s(f, B); put_u32s(f, dR); put_u32s(f, dC);
  std::fclose(f);
  return 0;
}
