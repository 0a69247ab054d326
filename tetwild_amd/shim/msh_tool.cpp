// msh_tool -- command-line access to msh_io (no GPU): converts between the
// .msh / .mesh files TetWild writes and a raw array file, so the reader and
// writer can be checked from outside.
//
//   msh_tool write  raw.bin out.msh [ascii]   raw -> .msh (binary by default)
//   msh_tool read   in.msh  raw.bin           .msh -> raw
//   msh_tool mwrite raw.bin out.mesh          raw -> Medit .mesh (tets)
//   msh_tool mread  in.mesh raw.bin           Medit .mesh -> raw
//
// raw.bin (little-endian): int64 nv, ne, k (nodes per element), type (msh
// element type), has_node_field, has_elem_field; double V[nv][3];
// int32 E[ne][k] (0-based); then double node_field[nv] "value" and
// double elem_field[ne] "min_dihedral_angle" when present.
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "msh_io.h"

using namespace tetwild_amd;

namespace {

struct Raw {
    int64_t nv = 0, ne = 0, k = 4, type = 4;
    std::vector<double> V, nf, ef;
    std::vector<int> E;
};

Raw read_raw(const char* path)
{
    Raw r;
    FILE* f = fopen(path, "rb");
    if (!f) throw std::runtime_error("cannot open raw input");
    int64_t h[6];
    if (fread(h, 8, 6, f) != 6) throw std::runtime_error("raw header");
    r.nv = h[0];
    r.ne = h[1];
    r.k = h[2];
    r.type = h[3];
    r.V.resize(3 * r.nv);
    std::vector<int32_t> e(r.ne * r.k);
    if (fread(r.V.data(), 8, r.V.size(), f) != r.V.size()) throw std::runtime_error("raw V");
    if (fread(e.data(), 4, e.size(), f) != e.size()) throw std::runtime_error("raw E");
    r.E.assign(e.begin(), e.end());
    if (h[4]) {
        r.nf.resize(r.nv);
        if (fread(r.nf.data(), 8, r.nv, f) != (size_t)r.nv) throw std::runtime_error("raw nf");
    }
    if (h[5]) {
        r.ef.resize(r.ne);
        if (fread(r.ef.data(), 8, r.ne, f) != (size_t)r.ne) throw std::runtime_error("raw ef");
    }
    fclose(f);
    return r;
}

void write_raw(const char* path, const Raw& r)
{
    FILE* f = fopen(path, "wb");
    if (!f) throw std::runtime_error("cannot open raw output");
    int64_t h[6] = {r.nv, r.ne, r.k, r.type, r.nf.empty() ? 0 : 1, r.ef.empty() ? 0 : 1};
    fwrite(h, 8, 6, f);
    fwrite(r.V.data(), 8, r.V.size(), f);
    std::vector<int32_t> e(r.E.begin(), r.E.end());
    fwrite(e.data(), 4, e.size(), f);
    if (!r.nf.empty()) fwrite(r.nf.data(), 8, r.nf.size(), f);
    if (!r.ef.empty()) fwrite(r.ef.data(), 8, r.ef.size(), f);
    fclose(f);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: msh_tool write|read|mwrite|mread <in> <out> [ascii]\n");
        return 2;
    }
    const std::string mode = argv[1];
    try {
        if (mode == "write") {
            Raw r = read_raw(argv[2]);
            MshSaver s(argv[3], !(argc > 4 && std::string(argv[4]) == "ascii"));
            s.save_mesh(r.V, r.E, 3, (MshSaver::ElementType)r.type);
            if (!r.nf.empty()) s.save_scalar_field("value", r.nf);
            if (!r.ef.empty()) s.save_elem_scalar_field("min_dihedral_angle", r.ef);
        } else if (mode == "read") {
            MshLoader l(argv[2]);
            Raw r;
            r.V = l.get_nodes();
            r.E = l.get_elements();
            r.nv = (int64_t)r.V.size() / 3;
            r.k = (int64_t)l.get_nodes_per_element();
            r.type = (int64_t)l.get_element_type();
            r.ne = r.k ? (int64_t)r.E.size() / r.k : 0;
            if (l.is_node_field("value")) r.nf = l.get_node_field("value");
            if (l.is_element_field("min_dihedral_angle"))
                r.ef = l.get_element_field("min_dihedral_angle");
            write_raw(argv[3], r);
        } else if (mode == "mwrite") {
            Raw r = read_raw(argv[2]);
            save_medit(argv[3], r.V, r.E);
        } else if (mode == "mread") {
            Raw r;
            load_medit(argv[2], r.V, r.E);
            r.nv = (int64_t)r.V.size() / 3;
            r.ne = (int64_t)r.E.size() / 4;
            write_raw(argv[3], r);
        } else {
            fprintf(stderr, "msh_tool: unknown mode %s\n", mode.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        fprintf(stderr, "msh_tool: %s\n", e.what());
        return 1;
    }
    return 0;
}
