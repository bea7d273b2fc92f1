// mesh_io.cpp -- see mesh_io.h.
#include "mesh_io.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace lats {

namespace {

// One corner token "i", "i/t", "i//n" or "i/t/n": the position index.
bool corner_index(const std::string& tok, long nverts, long& out, std::string& err) {
    const std::string head = tok.substr(0, tok.find('/'));
    if (head.empty()) {
        err = "face corner without a position index: '" + tok + "'";
        return false;
    }
    char* end = nullptr;
    errno = 0;
    const long v = std::strtol(head.c_str(), &end, 10);
    if (errno || *end != '\0' || v == 0) {
        err = "bad face index '" + tok + "'";
        return false;
    }
    out = v > 0 ? v - 1 : nverts + v;   // negative: relative to the vertices read so far
    if (out < 0) {
        err = "face index out of range: '" + tok + "'";
        return false;
    }
    return true;
}

}  // namespace

bool parse_obj(const std::string& text, ObjMesh& out, std::string& err) {
    out = ObjMesh();
    std::istringstream in(text);
    std::string line;
    size_t lineno = 0;
    std::vector<long> corners;
    while (std::getline(in, line)) {
        lineno++;
        // a trailing backslash continues the line
        while (!line.empty() && line.back() == '\\') {
            std::string next;
            if (!std::getline(in, next)) break;
            lineno++;
            line.pop_back();
            line += " " + next;
        }
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            double p[3];
            for (double& c : p)
                if (!(ls >> c)) {
                    err = "line " + std::to_string(lineno) + ": vertex needs three coordinates";
                    return false;
                }
            for (double c : p) out.x.push_back((float)c);
        } else if (tag == "f") {
            corners.clear();
            std::string tok;
            const long nv = (long)(out.x.size() / 3);
            while (ls >> tok) {
                long id;
                if (!corner_index(tok, nv, id, err)) {
                    err = "line " + std::to_string(lineno) + ": " + err;
                    return false;
                }
                corners.push_back(id);
            }
            if (corners.size() < 3) {
                err = "line " + std::to_string(lineno) + ": face with fewer than three corners";
                return false;
            }
            for (size_t k = 1; k + 1 < corners.size(); k++) {
                const long a = corners[0], b = corners[k], c = corners[k + 1];
                if (a == b || a == c || b == c) {
                    out.dropped++;
                    continue;
                }
                out.faces.push_back((int32_t)a);
                out.faces.push_back((int32_t)b);
                out.faces.push_back((int32_t)c);
            }
        }
        // vt, vn, o, g, s, usemtl, mtllib, l, p: nothing the cloth needs
    }
    const long nv = (long)(out.x.size() / 3);
    for (int32_t f : out.faces)
        if (f >= nv) {
            err = "face index " + std::to_string(f + 1) + " beyond the " + std::to_string(nv) + " vertices";
            return false;
        }
    return true;
}

bool read_obj(const std::string& path, ObjMesh& out, std::string& err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_obj(ss.str(), out, err);
}

std::string format_obj(const float* x, uint32_t nv, const int32_t* faces, uint32_t nf, uint32_t mesh_id) {
    std::ostringstream os;
    os << "# cloth frame written by libatssim_amd\n";
    os << "o mesh_" << mesh_id << "\n";
    for (uint32_t v = 0; v < nv; v++) os << "v " << x[3 * v] << " " << x[3 * v + 1] << " " << x[3 * v + 2] << "\n";
    for (uint32_t k = 0; k < nf; k++)
        os << "f " << faces[3 * k] + 1 << " " << faces[3 * k + 1] + 1 << " " << faces[3 * k + 2] + 1 << "\n";
    return os.str();
}

bool write_obj(const std::string& path, const float* x, uint32_t nv, const int32_t* faces, uint32_t nf,
               uint32_t mesh_id, std::string& err) {
    std::ofstream f(path, std::ios::out | std::ios::trunc);
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    f << format_obj(x, nv, faces, nf, mesh_id);
    if (!f) {
        err = "write failed: " + path;
        return false;
    }
    return true;
}

}  // namespace lats
