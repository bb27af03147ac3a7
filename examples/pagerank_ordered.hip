// pagerank_ordered.hip -- a non-idempotent functor set through the drop-in
// API: unnormalised PageRank in float, V(v) = 0.15 + 0.85 * sum over in-edges
// (u, v) of V(u) / outdeg(u), run with options::ordered (Jacobi iterations,
// a fixed combine order of each vertex's in-edges), so the result is defined
// bit for bit.  A vertex is updated while its value moves by more than 1e-4.
//
// usage: pagerank_ordered <edge-list> <out.f32>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "kites/kites.hpp"

static std::vector<unsigned> g_outdeg;

// Out-degrees from a first pass over the edge list (same line rules as the
// graph parser: a line whose first character is not a digit is skipped).
static void count_out_degrees(const char* path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error(std::string("Failed to open specified file: ") + path);
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] < '0' || line[0] > '9') continue;
    char* end = nullptr;
    unsigned long long const s = std::strtoull(line.c_str(), &end, 10);
    if (end == line.c_str() || (*end != ' ' && *end != '\t')) continue;
    if (g_outdeg.size() <= s) g_outdeg.resize(s + 1, 0);
    ++g_outdeg[s];
  }
}

template <typename vT, class eT>
bool pre_processing_pagerank(const unsigned, char**, const unsigned src, const unsigned,
                             vT& src_vertex, vT& dst_vertex, eT& edge) {
  src_vertex = 1.0f;
  dst_vertex = 1.0f;
  edge = 1.0f / static_cast<float>(g_outdeg[src]);
  return true;
}

int main(int argc, char** argv) {
  try {
    if (argc < 3) {
      std::cerr << "usage: " << argv[0] << " <edge-list> <out.f32>\n";
      return EXIT_FAILURE;
    }
    using vT = float;
    using eT = float;
    count_out_degrees(argv[1]);
    auto initF = [] __host__ __device__(volatile vT& locV, vT&) { locV = 0.15f; };
    auto nbrCompF = [] __host__ __device__(volatile vT& partialV, vT srcV, eT* connE) {
      partialV = 0.85f * srcV * (*connE);
    };
    auto redF = [] __host__ __device__(volatile vT& lhsV, volatile vT& rhsV) {
      lhsV = lhsV + rhsV;
    };
    auto updF = [] __host__ __device__(volatile vT& newV, vT oldV) {
      float const d = newV - oldV;
      return d > 1e-4f || d < -1e-4f;
    };
    kites::graph_csr<vT, eT> hostGrph(argv[1], pre_processing_pagerank<vT, eT>, 1,
                                      kites::input_graph_form::edge_list_s_d);
    kites::nv_gpu mydev(0);
    auto devGrph = kites::make_graph_csr_for_device(hostGrph, mydev);
    kites::options o;
    o.ordered = true;
    o.verbose = true;
    kites::process<kites::launch::sync>(devGrph, mydev, initF, nbrCompF, redF, updF, o);
    kites::host_pinned_buffer<vT> out(devGrph.V);
    std::ofstream f(argv[2], std::ios::binary);
    f.write(reinterpret_cast<const char*>(out.get_ptr()), out.sizeInBytes());
    std::cout << "Done." << std::endl;
    return EXIT_SUCCESS;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\nExiting." << std::endl;
    return EXIT_FAILURE;
  }
}
