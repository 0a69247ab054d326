// msh_io.cpp -- Gmsh 2.2 .msh (PyMesh MshSaver / MshLoader layout) and Medit
// .mesh files, as TetWild writes them (src/main.cpp:27-70).
#include "msh_io.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace tetwild_amd {

namespace {

int nodes_per_type(int type)
{
    switch (type) {
    case MshSaver::TRI: return 3;
    case MshSaver::QUAD: return 4;
    case MshSaver::TET: return 4;
    case MshSaver::HEX: return 8;
    default: throw std::runtime_error("msh: unsupported element type " + std::to_string(type));
    }
}

template <class T>
void put(std::ofstream& o, T v)
{
    o.write(reinterpret_cast<const char*>(&v), sizeof v);
}

template <class T>
T get(std::ifstream& in)
{
    T v;
    in.read(reinterpret_cast<char*>(&v), sizeof v);
    if (!in) throw std::runtime_error("msh: truncated binary block");
    return v;
}

// after a count line: spaces / tabs / CR, then the one newline that ends it
void end_of_line(std::ifstream& in)
{
    for (;;) {
        int c = in.get();
        if (c == '\n' || c == EOF) return;
        if (c != ' ' && c != '\t' && c != '\r')
            throw std::runtime_error("msh: unexpected data after a count");
    }
}

void expect(std::ifstream& in, const char* word)
{
    std::string w;
    in >> w;
    if (w != word) throw std::runtime_error(std::string("msh: expected ") + word + ", got " + w);
}

}  // namespace

// ---- writer -----------------------------------------------------------------
MshSaver::MshSaver(const std::string& filename, bool binary) : binary_(binary)
{
    out_.open(filename, binary ? std::ios::out | std::ios::binary : std::ios::out);
    if (!out_) throw std::runtime_error("msh: cannot open " + filename + " for writing");
}

MshSaver::~MshSaver() { out_.close(); }

void MshSaver::save_mesh(const std::vector<double>& nodes, const std::vector<int>& elements,
                         size_t dim, ElementType type)
{
    if (dim != 2 && dim != 3) throw std::runtime_error("msh: dimension must be 2 or 3");
    dim_ = dim;
    save_header();
    save_nodes(nodes);
    save_elements(elements, type);
}

void MshSaver::save_header()
{
    out_ << "$MeshFormat\n";
    if (binary_) {
        out_ << "2.2 1 " << sizeof(double) << "\n";
        put<int32_t>(out_, 1);  // endianness check, directly followed by the end tag
    } else {
        out_ << "2.2 0 " << sizeof(double) << "\n";
    }
    out_ << "$EndMeshFormat\n";
}

void MshSaver::save_nodes(const std::vector<double>& nodes)
{
    if (dim_ == 0) dim_ = 3;
    num_nodes_ = nodes.size() / dim_;
    out_ << "$Nodes\n" << num_nodes_ << "\n";
    for (size_t i = 0; i < num_nodes_; ++i) {
        const double* v = nodes.data() + i * dim_;
        const double z = dim_ == 3 ? v[2] : 0.0;
        if (binary_) {
            put<int32_t>(out_, (int32_t)(i + 1));
            put(out_, v[0]);
            put(out_, v[1]);
            put(out_, z);
        } else {
            out_ << i + 1 << " " << v[0] << " " << v[1] << " " << z << "\n";
        }
    }
    out_ << "$EndNodes\n";
}

void MshSaver::save_elements(const std::vector<int>& elements, ElementType type)
{
    const size_t k = (size_t)nodes_per_type(type);
    num_elements_ = elements.size() / k;
    out_ << "$Elements\n" << num_elements_ << "\n";
    if (num_elements_ > 0) {
        if (binary_) {
            put<int32_t>(out_, (int32_t)type);
            put<int32_t>(out_, (int32_t)num_elements_);
            put<int32_t>(out_, 0);  // no tags
        }
        for (size_t i = 0; i < num_elements_; ++i) {
            if (binary_) {
                put<int32_t>(out_, (int32_t)(i + 1));
                for (size_t j = 0; j < k; ++j) put<int32_t>(out_, elements[i * k + j] + 1);
            } else {
                out_ << i + 1 << " " << (int)type << " " << 0 << " ";
                for (size_t j = 0; j < k; ++j) out_ << elements[i * k + j] + 1 << " ";
                out_ << "\n";
            }
        }
    }
    out_ << "$EndElements\n";
}

void MshSaver::field(const char* section, const std::string& name, int components, size_t count,
                     const std::vector<double>& values, size_t stride)
{
    out_ << "$" << section << "\n1\n\"" << name << "\"\n1\n0.0\n3\n0\n" << components << "\n"
         << count << "\n";
    for (size_t i = 0; i < count; ++i) {
        double c[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < components && j < (int)stride; ++j) c[j] = values[i * stride + j];
        if (binary_) {
            put<int32_t>(out_, (int32_t)(i + 1));
            for (int j = 0; j < components; ++j) put(out_, c[j]);
        } else {
            out_ << i + 1;
            for (int j = 0; j < components; ++j) out_ << " " << c[j];
            out_ << "\n";
        }
    }
    out_ << "$End" << section << "\n";
}

void MshSaver::save_scalar_field(const std::string& name, const std::vector<double>& f)
{
    if (f.size() != num_nodes_) throw std::runtime_error("msh: node field size");
    field("NodeData", name, 1, num_nodes_, f, 1);
}

void MshSaver::save_vector_field(const std::string& name, const std::vector<double>& f)
{
    if (f.size() != num_nodes_ * dim_) throw std::runtime_error("msh: node field size");
    field("NodeData", name, 3, num_nodes_, f, dim_);
}

void MshSaver::save_elem_scalar_field(const std::string& name, const std::vector<double>& f)
{
    if (f.size() != num_elements_) throw std::runtime_error("msh: element field size");
    field("ElementData", name, 1, num_elements_, f, 1);
}

void MshSaver::save_elem_vector_field(const std::string& name, const std::vector<double>& f)
{
    if (f.size() != num_elements_ * dim_) throw std::runtime_error("msh: element field size");
    field("ElementData", name, 3, num_elements_, f, dim_);
}

// ---- reader -----------------------------------------------------------------
MshLoader::MshLoader(const std::string& filename)
{
    std::ifstream in(filename, std::ios::in | std::ios::binary);
    if (!in) throw std::runtime_error("msh: cannot open " + filename);
    expect(in, "$MeshFormat");
    double version;
    int type;
    size_t data_size;
    in >> version >> type >> data_size;
    if (!in || data_size != 8) throw std::runtime_error("msh: only 8-byte reals are supported");
    binary_ = type == 1;
    if (binary_) {
        end_of_line(in);
        if (get<int32_t>(in) != 1) throw std::runtime_error("msh: other-endian binary file");
    }
    expect(in, "$EndMeshFormat");
    std::string tag;
    while (in >> tag) {
        if (tag == "$Nodes") {
            nodes(in);
            expect(in, "$EndNodes");
        } else if (tag == "$Elements") {
            elements(in);
            expect(in, "$EndElements");
        } else if (tag == "$NodeData") {
            field(in, true);
            expect(in, "$EndNodeData");
        } else if (tag == "$ElementData") {
            field(in, false);
            expect(in, "$EndElementData");
        } else if (tag.size() > 1 && tag[0] == '$') {
            skip_section(in, tag.substr(1));
        } else {
            throw std::runtime_error("msh: unexpected token " + tag);
        }
    }
}

void MshLoader::nodes(std::ifstream& in)
{
    size_t n;
    in >> n;
    nodes_.assign(3 * n, 0.0);
    if (binary_) end_of_line(in);
    for (size_t i = 0; i < n; ++i) {
        int64_t idx;
        double v[3];
        if (binary_) {
            idx = get<int32_t>(in);
            for (double& c : v) c = get<double>(in);
        } else {
            in >> idx >> v[0] >> v[1] >> v[2];
        }
        if (idx < 1 || (size_t)idx > n) throw std::runtime_error("msh: node index out of range");
        for (int c = 0; c < 3; ++c) {
            if (!std::isfinite(v[c])) throw std::runtime_error("msh: NaN or Inf in the nodes");
            nodes_[3 * (idx - 1) + c] = v[c];
        }
    }
}

void MshLoader::elements(std::ifstream& in)
{
    size_t n;
    in >> n;
    std::map<int, std::vector<int>> by_type;
    if (binary_) {
        end_of_line(in);
        size_t done = 0;
        while (done < n) {
            const int type = get<int32_t>(in), count = get<int32_t>(in), tags = get<int32_t>(in);
            const int k = nodes_per_type(type);
            std::vector<int>& dst = by_type[type];
            for (int i = 0; i < count; ++i) {
                (void)get<int32_t>(in);  // element number
                for (int t = 0; t < tags; ++t) (void)get<int32_t>(in);
                for (int j = 0; j < k; ++j) dst.push_back(get<int32_t>(in) - 1);
            }
            done += (size_t)count;
        }
    } else {
        for (size_t i = 0; i < n; ++i) {
            int num, type, tags, t;
            in >> num >> type >> tags;
            for (int j = 0; j < tags; ++j) in >> t;
            const int k = nodes_per_type(type);
            std::vector<int>& dst = by_type[type];
            for (int j = 0; j < k; ++j) {
                in >> t;
                dst.push_back(t - 1);
            }
        }
    }
    // one element kind is kept, tets first (as PyMesh's loader does)
    for (int type : {(int)MshSaver::TET, (int)MshSaver::HEX, (int)MshSaver::TRI, (int)MshSaver::QUAD}) {
        auto it = by_type.find(type);
        if (it != by_type.end() && !it->second.empty()) {
            elements_ = it->second;
            element_type_ = (size_t)type;
            nodes_per_element_ = (size_t)nodes_per_type(type);
            return;
        }
    }
    element_type_ = 2;  // no elements: triangles by default
    nodes_per_element_ = 3;
}

void MshLoader::field(std::ifstream& in, bool per_node)
{
    size_t n_str, n_real, n_int;
    in >> n_str;
    std::vector<std::string> str(n_str);
    for (auto& s : str) {
        in >> std::ws;
        if (in.peek() == '"') {
            in.get();
            std::getline(in, s, '"');
        } else {
            in >> s;
        }
    }
    in >> n_real;
    for (size_t i = 0; i < n_real; ++i) {
        double r;
        in >> r;
    }
    in >> n_int;
    std::vector<long long> ints(n_int);
    for (auto& v : ints) in >> v;
    if (n_str < 1 || n_int < 3) throw std::runtime_error("msh: malformed field header");
    const int comps = (int)ints[1];
    const size_t count = (size_t)ints[2];
    std::vector<double>& f = per_node ? node_fields_[str[0]] : elem_fields_[str[0]];
    f.assign(count * comps, 0.0);
    if (binary_) end_of_line(in);
    for (size_t i = 0; i < count; ++i) {
        int64_t idx;
        if (binary_) idx = get<int32_t>(in);
        else in >> idx;
        if (idx < 1 || (size_t)idx > count) throw std::runtime_error("msh: field index out of range");
        for (int j = 0; j < comps; ++j) {
            double v;
            if (binary_) v = get<double>(in);
            else in >> v;
            f[(idx - 1) * comps + j] = v;
        }
    }
}

void MshLoader::skip_section(std::ifstream& in, const std::string& name)
{
    const std::string end = "$End" + name;
    std::string line;
    while (std::getline(in, line))
        if (line.compare(0, end.size(), end) == 0) return;
    throw std::runtime_error("msh: unterminated section " + name);
}

std::vector<std::string> MshLoader::get_node_field_names() const
{
    std::vector<std::string> r;
    for (const auto& kv : node_fields_) r.push_back(kv.first);
    return r;
}

std::vector<std::string> MshLoader::get_element_field_names() const
{
    std::vector<std::string> r;
    for (const auto& kv : elem_fields_) r.push_back(kv.first);
    return r;
}

// ---- Medit ------------------------------------------------------------------
void save_medit(const std::string& filename, const std::vector<double>& V, const std::vector<int>& T)
{
    std::ofstream f(filename);
    if (!f) throw std::runtime_error("mesh: cannot open " + filename + " for writing");
    f.precision(std::numeric_limits<double>::digits10 + 1);
    const size_t nv = V.size() / 3, nt = T.size() / 4;
    f << "MeshVersionFormatted 1\nDimension 3\nVertices\n" << nv << "\n";
    for (size_t i = 0; i < nv; ++i)
        f << V[3 * i] << " " << V[3 * i + 1] << " " << V[3 * i + 2] << " " << 0 << "\n";
    f << "Triangles\n" << 0 << "\nTetrahedra\n" << nt << "\n";
    for (size_t i = 0; i < nt; ++i) {
        for (int j = 0; j < 4; ++j) f << T[4 * i + j] + 1 << " ";
        f << 0 << "\n";
    }
    f << "End";
}

void load_medit(const std::string& filename, std::vector<double>& V, std::vector<int>& T)
{
    std::ifstream f(filename);
    if (!f) throw std::runtime_error("mesh: cannot open " + filename);
    V.clear();
    T.clear();
    std::string w;
    while (f >> w) {
        if (w == "Vertices") {
            size_t n;
            f >> n;
            V.resize(3 * n);
            for (size_t i = 0; i < n; ++i) {
                int ref;
                f >> V[3 * i] >> V[3 * i + 1] >> V[3 * i + 2] >> ref;
            }
        } else if (w == "Triangles") {
            size_t n;
            f >> n;
            for (size_t i = 0; i < 4 * n; ++i) {
                int skip;
                f >> skip;
            }
        } else if (w == "Tetrahedra") {
            size_t n;
            f >> n;
            T.resize(4 * n);
            for (size_t i = 0; i < n; ++i) {
                int ref;
                for (int j = 0; j < 4; ++j) {
                    f >> T[4 * i + j];
                    T[4 * i + j] -= 1;
                }
                f >> ref;
            }
        } else if (w == "End") {
            break;
        }
    }
    if (!f && !f.eof()) throw std::runtime_error("mesh: malformed file " + filename);
}

}  // namespace tetwild_amd
