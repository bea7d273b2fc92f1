// mesh_io.h -- Wavefront OBJ input and output of triangle cloths.
//
// Reading follows what the reference's SimMesh::read_mesh_file
// (src/Utils/mesh_reader.cpp:11-200, tinyobjloader with its default
// triangulation) keeps of a file: the `v` positions in file order, and every
// face fan-triangulated from its first corner ((0,1,2), (0,2,3), ...), with
// 1-based or negative (relative) indices and the `v/vt/vn` corner forms;
// a triangle that repeats a vertex is dropped (mesh_reader.cpp:172-182).
// Writing follows SolverInterface::save_mesh_to_obj
// (examples/example2_async_iter.cpp:1649-1710): a comment line, `o mesh_<k>`,
// `v x y z` with the stream's default 6 significant digits, `f a b c` 1-based.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lats {

struct ObjMesh {
    std::vector<float> x;         // 3 per vertex
    std::vector<int32_t> faces;   // 3 per triangle, 0-based
    uint32_t dropped = 0;         // triangles dropped for a repeated vertex
};

// Parses OBJ text; false with `err` set on malformed input.
bool parse_obj(const std::string& text, ObjMesh& out, std::string& err);
bool read_obj(const std::string& path, ObjMesh& out, std::string& err);

std::string format_obj(const float* x, uint32_t nv, const int32_t* faces, uint32_t nf, uint32_t mesh_id);
bool write_obj(const std::string& path, const float* x, uint32_t nv, const int32_t* faces, uint32_t nf,
               uint32_t mesh_id, std::string& err);

}  // namespace lats
