// Host-side setup for the Tet4 elastostatics path: box mesh generation, mesh sets,
// the node-node block graph and the LDS chunk partition used by the assembly kernels.
//
// Everything here runs once per mesh (the reference rebuilds the graph on every
// gradient_u call, src/ImplicitFunctors.hpp:735-808; here it is built once and kept).
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace pa {

// Structured box -> 6 Tet4 per hex, numbered like Omega_h::build_box
// (unit_tests/PlatoTestHelpers.hpp:206-250 calls it): lexicographic hexes split
// along the v0-v6 body diagonal, vertices renumbered along a Hilbert curve.
// Element order: stable sort by smallest vertex number.  The reference's element
// order is not reproduced; residual, Jacobian, graph, compliance and gradients do not
// depend on it (only per-cell worksets do).
struct BoxMesh {
  int num_nodes = 0;
  int num_elems = 0;
  std::vector<double> coords;   // [3*num_nodes], x,y,z interleaved
  std::vector<int32_t> conn;    // [4*num_elems]
};
BoxMesh build_box(int nx, int ny, int nz, double lx, double ly, double lz);

// Hilbert curve permutation (new -> old) of a 3-D point set.
std::vector<int32_t> hilbert_order(const double* coords, int n);

// Nodes whose coordinate `axis` equals `value` within `tol`.
std::vector<int32_t> nodes_on_plane(const double* coords, int nn, int axis, double value, double tol);

// Boundary faces whose three vertices lie on the plane; returns face vertex triples
// (in the element's local order) and the owning element.
void faces_on_plane(const int32_t* conn, int ne, const double* coords, int axis, double value,
                    double tol, std::vector<int32_t>& face_verts, std::vector<int32_t>& face_elems);

// Block CRS graph of the node star plus the diagonal (CreateBlockMatrix,
// src/ImplicitFunctors.hpp:735-808). Columns ascending.
struct Graph {
  int num_rows = 0;
  std::vector<int32_t> rowmap;  // [num_rows+1]
  std::vector<int32_t> colidx;  // [nnz_b]
  int max_row_len = 0;
};
Graph build_graph(const int32_t* conn, int ne, int nn);
// Validate a caller-supplied graph against the mesh; fills max_row_len. Empty string on success.
std::string check_graph(Graph& g, const int32_t* conn, int ne, int nn);

// node -> (element, local vertex) adjacency
struct NodeElems {
  std::vector<int32_t> offs;  // [nn+1]
  std::vector<int32_t> ent;   // element*4 + local
};
NodeElems node_elems(const int32_t* conn, int ne, int nn);

// Per-element slot codes: position of conn[b] in the row of conn[a], packed with
// `bits` bits per (a,b) pair (index 4a+b). bits = 4 (rows <= 16) or 8 (rows <= 256).
std::vector<uint32_t> slot_codes(const Graph& g, const int32_t* conn, int ne, int bits, std::string& err);

// Row chunks for the LDS assembly: consecutive rows whose blocks, rows and node
// table fit the LDS budget.  Per chunk:
//   * node table: the chunk's rows first (local index = row - first row), then the
//     other nodes its rows couple to (the union of the rows' graph columns), sorted;
//   * entries: every element touching one of the chunk's rows, as
//       list     element id (low 28 bits) | mask of owned vertices (top 4 bits),
//       ent_ln   the element's 4 vertices as local node-table indices (uint16),
//       ent_code the element's slot codes (words_per_code uint32 each),
//     sorted by owned-vertex count and interleaved (see build_chunks).
constexpr uint16_t kNoSlot = 0xFFFF;

struct Chunks {
  int count = 0;
  std::vector<int32_t> row_begin;    // [count+1]
  std::vector<int32_t> list_offs;    // [count+1]
  std::vector<uint32_t> list;        // element | (mask << 28)
  std::vector<int32_t> node_offs;    // [count+1]
  std::vector<int32_t> nodes;        // node tables
  std::vector<uint16_t> ent_ln;      // 4 per entry
  std::vector<uint32_t> ent_code;    // words_per_code per entry
  // per graph block (row r, column c): chunk-local slot of block (c, r) when both rows
  // lie in the chunk and c < r, else kNoSlot
  std::vector<uint16_t> tslot;
  int max_rows = 0, max_blocks = 0, max_nodes = 0;
};
Chunks build_chunks(const Graph& g, const NodeElems& ne_adj, const int32_t* conn, const std::vector<uint32_t>& codes,
                    int words_per_code, int max_blocks, int max_rows, int max_nodes);

// Minimal thread pool helper for setup loops.
void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)>& body);

}  // namespace pa
