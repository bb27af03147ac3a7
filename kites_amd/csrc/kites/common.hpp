// common.hpp -- API enums, constants and the throw-on-error check.
//
// Enums keep the reference's names and values (kites/common/globals.cuh:13-55)
// so user code written against the reference compiles unchanged.  Errors
// from the HIP runtime throw std::runtime_error with file and line, as the
// reference's CUDAErrorCheck does (kites/common/cuda_error_check.cuh:9-18).
#ifndef KITES_COMMON_HPP_
#define KITES_COMMON_HPP_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>

namespace kites {

enum {
  // Threads per workgroup of every engine kernel (4 waves of 64 lanes).
  COMPILE_TIME_DETERMINED_BLOCK_SIZE = 256,
  WAVE_SIZE = 64,
  WAVE_SIZE_SHIFT = 6,
  // Vertex count padding of host graphs (graph_raw.h:173-182).
  VERTEX_PADDING = 128
};

enum class launch { sync = true, async = false };

enum class operation_mode { iterative = true, frontier_based = false };

enum class input_graph_form { edge_list_s_d = true, edge_list_d_s = false };

enum class graph_property { directed = true, undirected = false };

enum class edge_list_expression { complete = true, incomplete = false };

enum class semi_dynamic_load_balancer { ON = true, OFF = false };

inline void hip_assert(hipError_t err, const char* file, int line) {
  if (err != hipSuccess) {
    std::ostringstream os;
    os << "HIPAssert: " << hipGetErrorString(err) << " " << file << " " << line
       << "\n";
    throw std::runtime_error(os.str());
  }
}

}  // namespace kites

#define KITES_HIP_CHECK(err) ::kites::hip_assert((err), __FILE__, __LINE__)

#endif  // KITES_COMMON_HPP_
