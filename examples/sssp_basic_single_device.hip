// sssp_basic_single_device.hip -- the reference's example
// (examples/sssp_basic_single_device.cu) written against the drop-in API:
// same pre-processing callback, same four lambdas, same pipeline
// (graph_csr -> gpu -> make_graph_csr_for_device -> process -> sanity_check).
//
// usage: sssp_basic_single_device <edge-list> [source=30] [vpg=1] [out.u32]
// Exits non-zero if the GPU result differs from the host fixpoint.
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
      std::cerr << "usage: " << argv[0] << " <edge-list> [source] [vpg] [out.u32]\n";
      return EXIT_FAILURE;
    }
    if (argc > 2) g_source = static_cast<unsigned>(std::atoi(argv[2]));
    unsigned const vpg = argc > 3 ? static_cast<unsigned>(std::atoi(argv[3])) : 1u;
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

    kites::graph_csr<vT, eT> hostGrph(argv[1], pre_processing_function_sssp<vT, eT>, vpg,
                                      kites::input_graph_form::edge_list_s_d);
    kites::nv_gpu mydev(0);
    std::cout << "Selected device: " << mydev.getDevIdx() << '\n';
    auto devGrph = kites::make_graph_csr_for_device(hostGrph, mydev);
    kites::process<kites::launch::sync>(devGrph, mydev, initF, nbrCompF, redF, updF);
    if (argc > 4) {
      kites::host_pinned_buffer<vT> out(devGrph.V);
      std::ofstream f(argv[4], std::ios::binary);
      f.write(reinterpret_cast<const char*>(out.get_ptr()), out.sizeInBytes());
    }
    bool const ok = kites::sanity_check(hostGrph, devGrph, initF, nbrCompF, redF, updF);
    std::cout << "Done." << std::endl;
    return ok ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\nExiting." << std::endl;
    return EXIT_FAILURE;
  }
}
