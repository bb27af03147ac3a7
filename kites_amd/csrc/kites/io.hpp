// io.hpp -- file access and text output of host graphs.
//
// openFileToAccess throws on failure (kites/io/file_interactor.hpp:15-20);
// out_vertices / save_to_file write the reference's text "prepared graph"
// format (kites/io/graph_output.cuh:11-38), which graph_csr(file) reads back.
#ifndef KITES_IO_HPP_
#define KITES_IO_HPP_

#include <fstream>
#include <stdexcept>
#include <string>

namespace kites {
namespace io {

template <typename T_file>
void openFileToAccess(T_file& f, const std::string& name) {
  f.open(name.c_str());
  if (!f) throw std::runtime_error("Failed to open specified file: " + name + "\n");
}

template <class G>
void out_vertices(G& g, std::ofstream& out) {
  out << g.V;
}
template <class G>
void out_vertices(G& g, std::string const& name) {
  std::ofstream out;
  openFileToAccess(out, name);
  out_vertices(g, out);
}

template <class G>
void save_to_file(G& g, std::ofstream& out) {
  out << g.get_num_vertices() << "\n" << g.get_num_edges() << "\n"
      << (g.gProp == decltype(g.gProp)::directed ? 1 : 0) << "\n";
  out << g.V << g.R << g.C << g.E << g.vBitmap;
  if (g.gProp == decltype(g.gProp)::directed) out << g.dir_R << g.dir_C;
}
template <class G>
void save_to_file(G& g, std::string const& name) {
  std::ofstream out;
  openFileToAccess(out, name);
  save_to_file(g, out);
}

}  // namespace io
}  // namespace kites

#endif  // KITES_IO_HPP_
