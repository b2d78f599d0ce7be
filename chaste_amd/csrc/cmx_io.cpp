// On-disk inputs of the monodomain path: Triangle/TetGen meshes (.node/.ele)
// and fibre orientation files (.ortho/.axi), read the way the reference's
// TrianglesMeshReader (mesh/src/reader/TrianglesMeshReader.cpp) and
// FibreReader (heart/src/io/FibreReader.cpp) read them, and the per-element
// conductivity tensors built from the fibre directions
// (heart/src/tissue/OrthotropicConductivityTensors.cpp,
// AxisymmetricConductivityTensors.cpp).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "cmx.h"

namespace {

thread_local std::string g_io_error;

int io_fail(const std::string& msg)
{
    g_io_error = msg;
    return CMX_E_ARG;
}

// Next line that is not blank once a '#' comment is removed; false at EOF.
bool next_data_line(std::istream& in, std::string& line)
{
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        if (line.find_first_not_of(" \t\r") != std::string::npos) return true;
    }
    return false;
}

struct MeshFile {
    int dim = 0, n_nodes = 0, n_node_attr = 0, n_bdy = 0;
    int n_elems = 0, nodes_per_elem = 0, n_elem_attr = 0;
    bool binary = false;
    std::vector<double> coords, elem_attr;
    std::vector<int32_t> elems;
};

// Header "n dim n_attributes n_boundary_markers [BIN]" of a .node file, and
// "n nodes_per_element n_attributes [BIN]" of a .ele file.
int read_mesh(const std::string& base, MeshFile& m)
{
    std::ifstream nf(base + ".node", std::ios::binary);
    if (!nf) return io_fail("cannot open " + base + ".node");
    std::string line;
    if (!next_data_line(nf, line)) return io_fail(base + ".node: no header");
    {
        std::istringstream h(line);
        std::string extra;
        if (!(h >> m.n_nodes >> m.dim >> m.n_node_attr >> m.n_bdy)) return io_fail(base + ".node: bad header");
        h >> extra;
        m.binary = extra == "BIN";
    }
    if (m.n_nodes <= 0 || m.dim < 1 || m.dim > 3 || m.n_node_attr < 0) return io_fail(base + ".node: bad header");
    m.coords.resize((size_t)m.n_nodes * m.dim);
    // ASCII files number their items from 0 or from 1 (the first item decides).
    int base_index = 0;
    if (m.binary) {
        // Binary items: dim doubles and the attributes as doubles, indexed from zero.
        std::vector<double> item(m.dim + m.n_node_attr);
        for (int i = 0; i < m.n_nodes; ++i) {
            if (!nf.read((char*)item.data(), item.size() * sizeof(double))) return io_fail(base + ".node: truncated");
            for (int d = 0; d < m.dim; ++d) m.coords[(size_t)i * m.dim + d] = item[d];
        }
    } else {
        for (int i = 0; i < m.n_nodes; ++i) {
            if (!next_data_line(nf, line)) return io_fail(base + ".node: unexpected end of file");
            std::istringstream s(line);
            long idx;
            if (!(s >> idx)) return io_fail(base + ".node: bad line");
            if (i == 0) {
                if (idx != 0 && idx != 1) return io_fail(base + ".node: first index must be 0 or 1");
                base_index = (int)idx;
            }
            if (idx != i + base_index) return io_fail(base + ".node: data for item " + std::to_string(i) + " missing");
            for (int d = 0; d < m.dim; ++d)
                if (!(s >> m.coords[(size_t)i * m.dim + d])) return io_fail(base + ".node: bad coordinate");
        }
    }
    std::ifstream ef(base + ".ele", std::ios::binary);
    if (!ef) return io_fail("cannot open " + base + ".ele");
    if (!next_data_line(ef, line)) return io_fail(base + ".ele: no header");
    bool ebin = false;
    {
        std::istringstream h(line);
        std::string extra;
        if (!(h >> m.n_elems >> m.nodes_per_elem >> m.n_elem_attr)) return io_fail(base + ".ele: bad header");
        h >> extra;
        ebin = extra == "BIN";
    }
    if (ebin != m.binary) return io_fail(base + ": .node and .ele disagree on BIN");
    if (m.n_elems <= 0 || m.nodes_per_elem != m.dim + 1 || m.n_elem_attr < 0)
        return io_fail(base + ".ele: only linear simplices of the node dimension are supported");
    m.elems.resize((size_t)m.n_elems * m.nodes_per_elem);
    m.elem_attr.resize((size_t)m.n_elems * m.n_elem_attr);
    for (int e = 0; e < m.n_elems; ++e) {
        int32_t* nd = &m.elems[(size_t)e * m.nodes_per_elem];
        double* at = m.n_elem_attr ? &m.elem_attr[(size_t)e * m.n_elem_attr] : nullptr;
        if (m.binary) {
            // node indices as 32-bit unsigned, then the attributes as doubles
            std::vector<uint32_t> idx(m.nodes_per_elem);
            if (!ef.read((char*)idx.data(), idx.size() * sizeof(uint32_t))) return io_fail(base + ".ele: truncated");
            for (int k = 0; k < m.nodes_per_elem; ++k) nd[k] = (int32_t)idx[k];
            if (m.n_elem_attr && !ef.read((char*)at, m.n_elem_attr * sizeof(double)))
                return io_fail(base + ".ele: truncated");
        } else {
            if (!next_data_line(ef, line)) return io_fail(base + ".ele: unexpected end of file");
            std::istringstream s(line);
            long idx;
            if (!(s >> idx) || idx != e + base_index)
                return io_fail(base + ".ele: data for item " + std::to_string(e) + " missing");
            for (int k = 0; k < m.nodes_per_elem; ++k) {
                long v;
                if (!(s >> v)) return io_fail(base + ".ele: bad node index");
                nd[k] = (int32_t)(v - base_index);
            }
            for (int a = 0; a < m.n_elem_attr; ++a)
                if (!(s >> at[a])) return io_fail(base + ".ele: error in reading attribute index " + std::to_string(a));
        }
        for (int k = 0; k < m.nodes_per_elem; ++k)
            if (nd[k] < 0 || nd[k] >= m.n_nodes) return io_fail(base + ".ele: node index out of range");
    }
    return CMX_OK;
}

}  // namespace

extern "C" {

const char* cmx_io_last_error(void) { return g_io_error.c_str(); }

int cmx_read_mesh(const char* base, int32_t* dim, int32_t* n_nodes, int32_t* n_elems, int32_t* n_elem_attr,
                  double* coords, int32_t* elems, double* elem_attr)
{
    if (!base || !dim || !n_nodes || !n_elems) return io_fail("null argument");
    MeshFile m;
    const int rc = read_mesh(base, m);
    if (rc) return rc;
    *dim = m.dim;
    *n_nodes = m.n_nodes;
    *n_elems = m.n_elems;
    if (n_elem_attr) *n_elem_attr = m.n_elem_attr;
    if (coords) std::memcpy(coords, m.coords.data(), m.coords.size() * sizeof(double));
    if (elems) std::memcpy(elems, m.elems.data(), m.elems.size() * sizeof(int32_t));
    if (elem_attr && !m.elem_attr.empty()) std::memcpy(elem_attr, m.elem_attr.data(), m.elem_attr.size() * sizeof(double));
    return CMX_OK;
}

int cmx_read_fibres(const char* path, int32_t per_item, int32_t* n_items, double* out)
{
    if (!path || !n_items || per_item <= 0) return io_fail("null argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return io_fail(std::string("Failed to open fibre file ") + path);
    std::string line;
    if (!next_data_line(f, line)) return io_fail(std::string(path) + ": no header");
    std::istringstream h(line);
    long n = -1;
    std::string extra, more;
    h >> n >> extra;
    const bool binary = extra == "BIN";
    if (n < 0 || (!extra.empty() && !binary) || (h >> more))
        return io_fail("First (non comment) line of the fibre orientation file should contain the number of "
                       "lines of data in the file (and possibly a BIN tag) at most");
    *n_items = (int32_t)n;
    if (!out) return CMX_OK;
    for (long e = 0; e < n; ++e) {
        double* row = out + (size_t)e * per_item;
        if (binary) {
            if (!f.read((char*)row, per_item * sizeof(double))) return io_fail(std::string(path) + ": truncated");
            continue;
        }
        if (!next_data_line(f, line)) return io_fail(std::string(path) + ": unexpected end of file");
        std::istringstream s(line);
        int got = 0;
        double v;
        while (s >> v) {
            if (got == per_item) return io_fail(std::string("Too many entries in a line in ") + path);
            row[got++] = v;
        }
        if (got < per_item) return io_fail(std::string("A line is incomplete in ") + path);
    }
    return CMX_OK;
}

int cmx_conductivity_tensors(int32_t dim, int32_t n_elems, int32_t axisymmetric, const double* fibres,
                             const double* sigma, double* out)
{
    if (dim < 1 || dim > 3 || n_elems < 0 || !fibres || !sigma || !out) return io_fail("bad argument");
    for (int e = 0; e < n_elems; ++e) {
        double* T = out + (size_t)e * dim * dim;
        if (axisymmetric) {
            // sigma_t I + (sigma_f - sigma_t) f f^T
            const double* f = fibres + (size_t)e * dim;
            for (int r = 0; r < dim; ++r)
                for (int c = 0; c < dim; ++c)
                    T[r * dim + c] = (r == c ? sigma[1] : 0.0) + (sigma[0] - sigma[1]) * f[r] * f[c];
        } else {
            // A diag(sigma) A^T with A's columns the fibre, sheet and normal
            // directions: the file's row k is direction k.
            const double* a = fibres + (size_t)e * dim * dim;
            for (int r = 0; r < dim; ++r)
                for (int c = 0; c < dim; ++c) {
                    double s = 0;
                    for (int k = 0; k < dim; ++k) s += a[k * dim + r] * sigma[k] * a[k * dim + c];
                    T[r * dim + c] = s;
                }
        }
    }
    return CMX_OK;
}

}  // extern "C"
