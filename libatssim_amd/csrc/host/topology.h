// topology.h -- cloth topology, rest state and vertex colouring (host, C++).
//
// Builds, from triangle positions + faces, every per-mesh array the VBD path
// consumes, in the reference's formats and orders:
//   * edges sorted by (min id, max id) and bending pairs
//     (e0, e1, opposite vertex in the lower-id face, opposite in the higher-id
//     face)                           -- mesh_reader.h:206-298
//   * vertex->edge and vertex->bending-pair lists in ascending element id,
//     as packed CSR (n+1 absolute offsets starting at n+1, then the data)
//                                     -- example2_async_iter.cpp:99-127, 229-261
//   * rest lengths and the quadratic-bending matrix Q = 3/(A0+A1) K K^T
//                                     -- example2_async_iter.cpp:308-374
//   * the greedy vertex colouring over edge + bending-pair neighbours and its
//     clusters (ascending vertex id)  -- example2_async_iter.cpp:572-646
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lats {

struct Topology {
    uint32_t num_verts = 0;
    std::vector<int32_t> faces;            // 3F
    std::vector<int32_t> edges;            // 2E
    std::vector<int32_t> bending;          // 4B
    std::vector<float> rest_length;        // E
    std::vector<float> bending_Q;          // 16B, [b][col][row]
    std::vector<uint32_t> adj_edges_csr;   // packed CSR, V rows
    std::vector<uint32_t> adj_bending_csr; // packed CSR, V rows
    std::vector<uint32_t> clusters_csr;    // packed CSR, C rows
    std::vector<uint8_t> colour;           // V, colour id of each vertex
    uint32_t num_clusters = 0;

    uint32_t num_edges() const { return (uint32_t)(edges.size() / 2); }
    uint32_t num_bending() const { return (uint32_t)(bending.size() / 4); }
};

// Builds everything above.  Faces with a repeated vertex must already be
// removed (the OBJ reader does so).  Returns false with `err` set on invalid
// input (index out of range, more than 255 colours).
bool build_topology(const float* x, uint32_t nv, const int32_t* faces, uint32_t nf, Topology& out,
                    std::string& err);

// The topology on disk: one raw array per "<dir>/<name>.dat" file (the
// reference's save_to_binary / read_from_binary format, mesh_reader.h:321-343).
// load_topology checks every array against the others.
bool save_topology(const Topology& t, const std::string& dir, std::string& err);
bool load_topology(const std::string& dir, Topology& out, std::string& err);

// Packed-CSR helpers: row i of a packed CSR is data[p[i]] .. data[p[i+1]-1].
inline uint32_t csr_rows(const std::vector<uint32_t>& p) { return p.empty() ? 0 : (uint32_t)(p[0] - 1); }

}  // namespace lats
