// topology_io.cpp -- the topology on disk, one raw array per file, the way
// the reference stores solver arrays (save_to_binary / read_from_binary,
// src/Utils/mesh_reader.h:321-343: the array's bytes as they sit in memory,
// "<name>.dat", the element count taken from the file size on reading), in a
// directory the caller names instead of the reference's hard-coded path.
// The adjacency and cluster arrays are packed CSRs as upload_2d_csr_from lays
// them out (examples/example2_async_iter.cpp:99-127); loading checks every
// array against the others before anything uses it.
#include <cstdio>
#include <string>
#include <vector>

#include "topology.h"

namespace lats {
namespace {

template <typename T> bool write_array(const std::string& dir, const char* name, const std::vector<T>& a,
                                       std::string& err) {
    const std::string path = dir + "/" + name + ".dat";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = "cannot open " + path + " for writing"; return false; }
    const size_t n = a.empty() ? 0 : std::fwrite(a.data(), sizeof(T), a.size(), f);
    const bool ok = n == a.size() && std::fclose(f) == 0;
    if (!ok) err = "cannot write " + path;
    return ok;
}

template <typename T> bool read_array(const std::string& dir, const char* name, std::vector<T>& a, std::string& err) {
    const std::string path = dir + "/" + name + ".dat";
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open " + path; return false; }
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long bytes = ok ? std::ftell(f) : -1;
    ok = ok && bytes >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok && bytes % (long)sizeof(T) != 0) {
        std::fclose(f);
        err = path + ": size is not a whole number of elements";
        return false;
    }
    if (ok) {
        a.resize((size_t)bytes / sizeof(T));
        ok = a.empty() || std::fread(a.data(), sizeof(T), a.size(), f) == a.size();
    }
    std::fclose(f);
    if (!ok) err = "cannot read " + path;
    return ok;
}

// A packed CSR of `rows` rows over ids < `bound` (rows == ~0u: any count).
bool check_csr(const std::vector<uint32_t>& p, uint32_t rows, uint32_t bound, const char* name, std::string& err) {
    if (p.empty() || p[0] == 0 || (rows != ~0u && p[0] != rows + 1) || p[0] > p.size()) {
        err = std::string(name) + ": packed CSR header does not match";
        return false;
    }
    const uint32_t n = p[0] - 1;
    if (p[n] != p.size()) { err = std::string(name) + ": last offset is not the array length"; return false; }
    for (uint32_t i = 0; i < n; i++)
        if (p[i] > p[i + 1]) { err = std::string(name) + ": offsets decrease"; return false; }
    for (size_t k = n + 1; k < p.size(); k++)
        if (p[k] >= bound) { err = std::string(name) + ": id out of range"; return false; }
    return true;
}

const char* kFiles[] = {"faces", "edges", "bending_edges", "edges_rest_state_length", "bending_edges_Q",
                        "vert_adj_edges", "vert_adj_bending_edges", "clusterd_per_vertex_bending",
                        "per_vertex_bending_cluster_id"};

}  // namespace

bool save_topology(const Topology& t, const std::string& dir, std::string& err) {
    return write_array(dir, kFiles[0], t.faces, err) && write_array(dir, kFiles[1], t.edges, err) &&
           write_array(dir, kFiles[2], t.bending, err) && write_array(dir, kFiles[3], t.rest_length, err) &&
           write_array(dir, kFiles[4], t.bending_Q, err) && write_array(dir, kFiles[5], t.adj_edges_csr, err) &&
           write_array(dir, kFiles[6], t.adj_bending_csr, err) && write_array(dir, kFiles[7], t.clusters_csr, err) &&
           write_array(dir, kFiles[8], t.colour, err);
}

bool load_topology(const std::string& dir, Topology& out, std::string& err) {
    Topology t;
    if (!(read_array(dir, kFiles[0], t.faces, err) && read_array(dir, kFiles[1], t.edges, err) &&
          read_array(dir, kFiles[2], t.bending, err) && read_array(dir, kFiles[3], t.rest_length, err) &&
          read_array(dir, kFiles[4], t.bending_Q, err) && read_array(dir, kFiles[5], t.adj_edges_csr, err) &&
          read_array(dir, kFiles[6], t.adj_bending_csr, err) && read_array(dir, kFiles[7], t.clusters_csr, err) &&
          read_array(dir, kFiles[8], t.colour, err)))
        return false;
    if (t.faces.size() % 3 || t.edges.size() % 2 || t.bending.size() % 4) {
        err = "faces, edges or bending pairs are not whole records";
        return false;
    }
    const uint32_t E = t.num_edges(), B = t.num_bending();
    if (t.rest_length.size() != E || t.bending_Q.size() != 16 * (size_t)B) {
        err = "rest lengths or Q do not match the edge / bending-pair counts";
        return false;
    }
    if (!check_csr(t.adj_edges_csr, ~0u, E, "vert_adj_edges", err)) return false;
    const uint32_t V = csr_rows(t.adj_edges_csr);
    if (!check_csr(t.adj_bending_csr, V, B, "vert_adj_bending_edges", err)) return false;
    if (!check_csr(t.clusters_csr, ~0u, V, "clusterd_per_vertex_bending", err)) return false;
    if (t.clusters_csr.size() - t.clusters_csr[0] != V) {
        err = "clusters do not hold every vertex once";
        return false;
    }
    if (t.colour.size() != V) { err = "per_vertex_bending_cluster_id has the wrong length"; return false; }
    for (int32_t v : t.faces)
        if (v < 0 || (uint32_t)v >= V) { err = "face vertex out of range"; return false; }
    for (int32_t v : t.edges)
        if (v < 0 || (uint32_t)v >= V) { err = "edge vertex out of range"; return false; }
    for (int32_t v : t.bending)
        if (v < 0 || (uint32_t)v >= V) { err = "bending vertex out of range"; return false; }
    t.num_verts = V;
    t.num_clusters = csr_rows(t.clusters_csr);
    for (uint32_t c = 0; c < t.num_clusters; c++)
        for (uint32_t r = t.clusters_csr[c]; r < t.clusters_csr[c + 1]; r++)
            if (t.colour[t.clusters_csr[r]] != c) {
                err = "per_vertex_bending_cluster_id disagrees with the clusters";
                return false;
            }
    out = std::move(t);
    return true;
}

}  // namespace lats
