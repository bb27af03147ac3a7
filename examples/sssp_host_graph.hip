// sssp_host_graph.hip -- the reference example's pipeline through the
// host-graph overload of process() (kites/device/compute.cuh:397-411):
// process<launch::sync>(hostGrph, dev, ...) uploads the graph, runs the
// fixpoint and copies the vertex values back into hostGrph.V, so the
// result is read from the host graph without a device graph in user code.
//
// usage: sssp_host_graph <edge-list> [source=30] [out.u32]
// The same callback and lambdas as examples/sssp_basic_single_device.hip.
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
    if (argc < 2) {
      std::cerr << "usage: " << argv[0] << " <edge-list> [source] [out.u32]\n";
      return EXIT_FAILURE;
    }
    if (argc > 2) g_source = static_cast<unsigned>(std::atoi(argv[2]));
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

    kites::graph_csr<vT, eT> hostGrph(argv[1], pre_processing_function_sssp<vT, eT>,
                                      kites::input_graph_form::edge_list_s_d);
    kites::nv_gpu mydev(0);
    std::cout << "Selected device: " << mydev.getDevIdx() << '\n';
    std::size_t const n = hostGrph.get_num_vertices();
    vT const before = n > g_source ? hostGrph.V[g_source] : 0u;
    kites::process<kites::launch::sync>(hostGrph, mydev, initF, nbrCompF, redF, updF);
    std::cout << "Source value before " << before << ", after " << hostGrph.V[g_source] << '\n';
    if (argc > 3) {
      std::ofstream f(argv[3], std::ios::binary);
      f.write(reinterpret_cast<const char*>(hostGrph.V.get_ptr()), hostGrph.V.sizeInBytes());
    }
    std::cout << "Done." << std::endl;
    return EXIT_SUCCESS;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\nExiting." << std::endl;
    return EXIT_FAILURE;
  }
}
