// prepared_graph.hip -- the prepared-graph path of the drop-in API: build a
// graph from an edge list (the reference example's callback), write it with
// kites::io::save_to_file (kites/io/graph_output.cuh:29-35), read it back with
// graph_csr(file) (kites/graph/graph_csr.h:112-166) and run the reference
// example's SSSP on the graph that was read.
//
// usage: prepared_graph <edge-list> <dump.txt> <out.u32> [source=30]
// Exits non-zero if the run on the reloaded graph fails the sanity check.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "kites/kites.hpp"

static unsigned g_source = 30;

template <typename vT, class eT>
bool pre_processing_function_sssp(const unsigned argcount, char** argvector,
                                  const unsigned src_vertex_index, const unsigned dst_vertex_index,
                                  vT& src_vertex, vT& dst_vertex, eT& edge) {
  src_vertex = (src_vertex_index != g_source) ? 1073741823 : 0;
  dst_vertex = (dst_vertex_index != g_source) ? 1073741823 : 0;
  edge = (argcount > 0) ? std::atoi(argvector[0])
                        : ((src_vertex_index + 1) * (dst_vertex_index + 2) % 127);
  return src_vertex_index == g_source;
}

int main(int argc, char** argv) {
  try {
    if (argc < 4) {
      std::cerr << "usage: " << argv[0] << " <edge-list> <dump.txt> <out.u32> [source]\n";
      return EXIT_FAILURE;
    }
    if (argc > 4) g_source = static_cast<unsigned>(std::atoi(argv[4]));
    using vT = unsigned int;
    using eT = unsigned int;
    auto initF = [] __host__ __device__(volatile vT& locV, vT& globV) { locV = globV; };
    auto nbrCompF = [] __host__ __device__(volatile vT& partialV, vT srcV, eT* connE) {
      partialV = srcV + (*connE);
    };
    auto redF = [] __host__ __device__(volatile vT& lhsV, volatile vT& rhsV) {
      if (lhsV > rhsV) lhsV = rhsV;
    };
    auto updF = [] __host__ __device__(volatile vT& newV, vT oldV) { return newV < oldV; };

    {
      kites::graph_csr<vT, eT> built(argv[1], pre_processing_function_sssp<vT, eT>, 1,
                                     kites::input_graph_form::edge_list_s_d);
      kites::io::save_to_file(built, std::string(argv[2]));
    }
    kites::graph_csr<vT, eT> hostGrph{std::string(argv[2])};
    kites::nv_gpu mydev(0);
    auto devGrph = kites::make_graph_csr_for_device(hostGrph, mydev);
    kites::process<kites::launch::sync>(devGrph, mydev, initF, nbrCompF, redF, updF);
    kites::host_pinned_buffer<vT> out(devGrph.V);
    std::ofstream f(argv[3], std::ios::binary);
    f.write(reinterpret_cast<const char*>(out.get_ptr()), out.sizeInBytes());
    bool const ok = kites::sanity_check(hostGrph, devGrph, initF, nbrCompF, redF, updF);
    std::cout << "Done." << std::endl;
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\nExiting." << std::endl;
    return EXIT_FAILURE;
  }
}
