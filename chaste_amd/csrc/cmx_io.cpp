// On-disk formats of the monodomain path: Triangle/TetGen meshes (.node/.ele)
// and fibre orientation files (.ortho/.axi), read the way the reference's
// TrianglesMeshReader (mesh/src/reader/TrianglesMeshReader.cpp) and
// FibreReader (heart/src/io/FibreReader.cpp) read them, and the per-element
// conductivity tensors built from the fibre directions
// (heart/src/tissue/OrthotropicConductivityTensors.cpp,
// AxisymmetricConductivityTensors.cpp); and the HDF5 results file (below).
#include <algorithm>
#include <cmath>
#include <cstdio>
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
int read_mesh(const std::string& base, MeshFile& m, int elem_dim = 0)
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
    // Elements of the node dimension come from <base>.ele ("n nodes_per_element
    // n_attributes [BIN]"); line elements of a mesh in 2-D or 3-D space
    // (TrianglesMeshReader<1, SPACE_DIM>) from <base>.edge, whose header is
    // "n n_attributes [BIN]" with two nodes per item.
    if (elem_dim <= 0) elem_dim = m.dim;
    if (elem_dim != m.dim && elem_dim != 1) return io_fail(base + ": only volume or line elements are supported");
    const std::string ext = elem_dim == m.dim ? ".ele" : ".edge";
    std::ifstream ef(base + ext, std::ios::binary);
    if (!ef) return io_fail("cannot open " + base + ext);
    if (!next_data_line(ef, line)) return io_fail(base + ext + ": no header");
    bool ebin = false;
    {
        std::istringstream h(line);
        std::string extra;
        if (elem_dim == m.dim) {
            if (!(h >> m.n_elems >> m.nodes_per_elem >> m.n_elem_attr)) return io_fail(base + ext + ": bad header");
        } else {
            if (!(h >> m.n_elems >> m.n_elem_attr)) return io_fail(base + ext + ": bad header");
            m.nodes_per_elem = 2;
        }
        h >> extra;
        ebin = extra == "BIN";
    }
    if (ebin != m.binary) return io_fail(base + ": .node and " + ext + " disagree on BIN");
    if (m.n_elems <= 0 || m.nodes_per_elem != elem_dim + 1 || m.n_elem_attr < 0)
        return io_fail(base + ext + ": only linear simplices are supported");
    m.elems.resize((size_t)m.n_elems * m.nodes_per_elem);
    m.elem_attr.resize((size_t)m.n_elems * m.n_elem_attr);
    for (int e = 0; e < m.n_elems; ++e) {
        int32_t* nd = &m.elems[(size_t)e * m.nodes_per_elem];
        double* at = m.n_elem_attr ? &m.elem_attr[(size_t)e * m.n_elem_attr] : nullptr;
        if (m.binary) {
            // node indices as 32-bit unsigned, then the attributes as doubles
            std::vector<uint32_t> idx(m.nodes_per_elem);
            if (!ef.read((char*)idx.data(), idx.size() * sizeof(uint32_t))) return io_fail(base + ext + ": truncated");
            for (int k = 0; k < m.nodes_per_elem; ++k) nd[k] = (int32_t)idx[k];
            if (m.n_elem_attr && !ef.read((char*)at, m.n_elem_attr * sizeof(double)))
                return io_fail(base + ext + ": truncated");
        } else {
            if (!next_data_line(ef, line)) return io_fail(base + ext + ": unexpected end of file");
            std::istringstream s(line);
            long idx;
            if (!(s >> idx) || idx != e + base_index)
                return io_fail(base + ext + ": data for item " + std::to_string(e) + " missing");
            for (int k = 0; k < m.nodes_per_elem; ++k) {
                long v;
                if (!(s >> v)) return io_fail(base + ext + ": bad node index");
                nd[k] = (int32_t)(v - base_index);
            }
            for (int a = 0; a < m.n_elem_attr; ++a)
                if (!(s >> at[a])) return io_fail(base + ext + ": error in reading attribute index " + std::to_string(a));
        }
        for (int k = 0; k < m.nodes_per_elem; ++k)
            if (nd[k] < 0 || nd[k] >= m.n_nodes) return io_fail(base + ext + ": node index out of range");
    }
    return CMX_OK;
}

}  // namespace

extern "C" {

const char* cmx_io_last_error(void) { return g_io_error.c_str(); }

int cmx_read_mesh(const char* base, int32_t* dim, int32_t* n_nodes, int32_t* n_elems, int32_t* n_elem_attr,
                  double* coords, int32_t* elems, double* elem_attr)
{
    return cmx_read_mesh_elements(base, 0, dim, n_nodes, n_elems, n_elem_attr, coords, elems, elem_attr);
}

int cmx_read_mesh_elements(const char* base, int32_t elem_dim, int32_t* dim, int32_t* n_nodes, int32_t* n_elems,
                           int32_t* n_elem_attr, double* coords, int32_t* elems, double* elem_attr)
{
    if (!base || !dim || !n_nodes || !n_elems) return io_fail("null argument");
    MeshFile m;
    const int rc = read_mesh(base, m, elem_dim);
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

// ---------------------------------------------------------------------------
// HDF5 results file in the reference's layout (Hdf5DataWriter,
// io/src/writer/Hdf5DataWriter.cpp:544-663): dataset "Data" [time][node][var]
// (IEEE double) with attributes "Variable Details" ("name(unit)" strings) and
// "IsDataComplete", dataset "Time" [time] with attribute "Unit".  Written from
// the HDF5 file format specification: version-0 superblock, a root group with
// a symbol table (one v1 B-tree leaf, one symbol node, a local heap), v1 object
// headers, contiguous storage.  The data is streamed to disk one time step at
// a time; the sizes in the headers and the time column are written at close.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kUndef = ~0ull;

struct Bytes {
    std::vector<uint8_t> b;
    size_t size() const { return b.size(); }
    void u8(uint32_t v) { b.push_back((uint8_t)v); }
    void u16(uint32_t v) { u8(v & 0xff); u8((v >> 8) & 0xff); }
    void u32(uint32_t v) { u16(v & 0xffff); u16(v >> 16); }
    void u64(uint64_t v) { u32((uint32_t)v); u32((uint32_t)(v >> 32)); }
    void raw(const void* p, size_t n) { b.insert(b.end(), (const uint8_t*)p, (const uint8_t*)p + n); }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void pad8() { while (b.size() % 8) u8(0); }
    void put64(size_t at, uint64_t v) { for (int k = 0; k < 8; ++k) b[at + k] = (uint8_t)(v >> (8 * k)); }
};

// Message bodies ------------------------------------------------------------
Bytes dataspace_msg(const std::vector<uint64_t>& dims, size_t* dims_at)
{
    Bytes m;
    m.u8(1);                 // version
    m.u8((uint32_t)dims.size());
    m.u8(0);                 // flags: no maximum dimensions
    m.u8(0);
    m.u32(0);                // reserved
    if (dims_at) *dims_at = m.size();
    for (uint64_t d : dims) m.u64(d);
    return m;
}

Bytes double_type()
{
    Bytes m;
    m.u8(0x11);              // version 1, class 1 (floating point)
    m.u8(0x20);              // little endian, mantissa normalisation: implied msb
    m.u8(63);                // sign bit position
    m.u8(0);
    m.u32(8);                // size
    m.u16(0);                // bit offset
    m.u16(64);               // precision
    m.u8(52);                // exponent location
    m.u8(11);                // exponent size
    m.u8(0);                 // mantissa location
    m.u8(52);                // mantissa size
    m.u32(1023);             // exponent bias
    return m;
}

Bytes string_type(uint32_t size)
{
    Bytes m;
    m.u8(0x13);              // version 1, class 3 (string)
    m.u8(0x00);              // null terminated, ASCII
    m.u8(0);
    m.u8(0);
    m.u32(size);
    return m;
}

Bytes uint32_type()
{
    Bytes m;
    m.u8(0x10);              // version 1, class 0 (fixed point)
    m.u8(0x00);              // little endian, unsigned
    m.u8(0);
    m.u8(0);
    m.u32(4);
    m.u16(0);
    m.u16(32);
    return m;
}

Bytes fill_msg()
{
    Bytes m;
    m.u8(2);                 // version 2
    m.u8(1);                 // space allocation: early
    m.u8(2);                 // fill value write time: never
    m.u8(0);                 // no fill value defined
    return m;
}

Bytes layout_msg(size_t* addr_at)
{
    Bytes m;
    m.u8(3);                 // version 3
    m.u8(1);                 // contiguous
    if (addr_at) *addr_at = m.size();
    m.u64(kUndef);           // address (patched)
    m.u64(0);                // size (patched)
    return m;
}

Bytes attribute_msg(const std::string& name, const Bytes& type, const Bytes& space, const Bytes& data)
{
    Bytes m;
    m.u8(1);                 // version 1
    m.u8(0);
    m.u16((uint32_t)name.size() + 1);
    m.u16((uint32_t)type.size());
    m.u16((uint32_t)space.size());
    m.raw(name.c_str(), name.size() + 1);
    m.pad8();
    m.raw(type.b.data(), type.size());
    m.pad8();
    m.raw(space.b.data(), space.size());
    m.pad8();
    m.raw(data.b.data(), data.size());
    return m;
}

// v1 object header holding the given messages; *at[k] gets the file offset
// of message k's body.
Bytes object_header(const std::vector<std::pair<uint16_t, Bytes>>& msgs, size_t file_at, std::vector<size_t>* at)
{
    Bytes body;
    for (const auto& mm : msgs) {
        Bytes m = mm.second;
        m.pad8();
        body.u16(mm.first);
        body.u16((uint32_t)m.size());
        body.u8(0);          // flags
        body.zeros(3);
        if (at) at->push_back(file_at + 16 + body.size());
        body.raw(m.b.data(), m.size());
    }
    Bytes h;
    h.u8(1);                 // version
    h.u8(0);
    h.u16((uint32_t)msgs.size());
    h.u32(1);                // reference count
    h.u32((uint32_t)body.size());
    h.u32(0);                // alignment to 8
    h.raw(body.b.data(), body.size());
    return h;
}

}  // namespace

struct cmx_h5 {
    std::FILE* f = nullptr;
    int32_t n_nodes = 0, n_vars = 0;
    uint64_t steps = 0;
    uint64_t data_addr = 0;
    std::vector<double> times;
    // patch points (file offsets)
    size_t eof_at = 0, data_dims_at = 0, data_layout_at = 0, time_dims_at = 0, time_layout_at = 0;
};

extern "C" {

int cmx_h5_open(const char* path, int32_t n_nodes, int32_t n_vars, const char* const* names,
                const char* const* units, const char* time_unit, cmx_h5** out)
{
    if (!path || !out || n_nodes <= 0 || n_vars <= 0 || !names || !units) return io_fail("bad argument");
    *out = nullptr;
    cmx_h5* h = new cmx_h5;
    h->n_nodes = n_nodes;
    h->n_vars = n_vars;
    h->f = std::fopen(path, "wb");
    if (!h->f) {
        delete h;
        return io_fail(std::string("cannot create ") + path);
    }
    // layout: superblock 96 | root header | local heap (+64 data) | B-tree
    // node | symbol node | Data header | Time header | raw data ...
    const uint64_t sb = 0, root = 96, root_size = 16 + 8 + 16, heap = root + root_size, heap_data = heap + 32,
                   btree = heap_data + 64, btree_size = 24 + 33 * 8 + 32 * 8, snod = btree + btree_size,
                   snod_size = 8 + 8 * 40, data_oh = snod + snod_size;
    (void)sb;
    // Data header
    std::vector<uint64_t> ddims = {0, (uint64_t)n_nodes, (uint64_t)n_vars};
    size_t d_dims_off = 0, d_lay_off = 0;
    Bytes dspace = dataspace_msg(ddims, &d_dims_off);
    Bytes dlay = layout_msg(&d_lay_off);
    std::vector<std::string> details;
    uint32_t smax = 1;
    for (int v = 0; v < n_vars; ++v) {
        details.push_back(std::string(names[v]) + "(" + units[v] + ")");
        smax = std::max<uint32_t>(smax, (uint32_t)details.back().size() + 1);
    }
    Bytes vd_data;
    for (const auto& s : details) {
        vd_data.raw(s.c_str(), s.size());
        vd_data.zeros(smax - s.size());
    }
    Bytes one;
    one.u32(1);
    std::vector<std::pair<uint16_t, Bytes>> dmsgs = {
        {0x01, dspace}, {0x03, double_type()}, {0x05, fill_msg()}, {0x08, dlay},
        {0x0C, attribute_msg("Variable Details", string_type(smax), dataspace_msg({(uint64_t)n_vars}, nullptr), vd_data)},
        {0x0C, attribute_msg("IsDataComplete", uint32_type(), dataspace_msg({1}, nullptr), one)}};
    std::vector<size_t> dat;
    Bytes dhdr = object_header(dmsgs, data_oh, &dat);
    const uint64_t time_oh = data_oh + dhdr.size();
    size_t t_dims_off = 0, t_lay_off = 0;
    Bytes tspace = dataspace_msg({0}, &t_dims_off);
    Bytes tlay = layout_msg(&t_lay_off);
    const std::string tu = time_unit ? time_unit : "msecs";
    Bytes tu_data;
    tu_data.raw(tu.c_str(), tu.size() + 1);
    std::vector<std::pair<uint16_t, Bytes>> tmsgs = {
        {0x01, tspace}, {0x03, double_type()}, {0x05, fill_msg()}, {0x08, tlay},
        {0x0C, attribute_msg("Unit", string_type((uint32_t)tu.size() + 1), dataspace_msg({1}, nullptr), tu_data)}};
    std::vector<size_t> tat;
    Bytes thdr = object_header(tmsgs, time_oh, &tat);
    h->data_addr = (time_oh + thdr.size() + 7) / 8 * 8;
    h->data_dims_at = dat[0] + d_dims_off;
    h->data_layout_at = dat[3] + d_lay_off;
    h->time_dims_at = tat[0] + t_dims_off;
    h->time_layout_at = tat[3] + t_lay_off;

    Bytes file;
    // superblock
    const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    file.raw(sig, 8);
    file.u8(0); file.u8(0); file.u8(0); file.u8(0);   // versions, reserved
    file.u8(0); file.u8(8); file.u8(8); file.u8(0);   // shared header version, offset/length sizes
    file.u16(4);                                       // group leaf node K
    file.u16(16);                                      // group internal node K
    file.u32(0);                                       // consistency flags
    file.u64(0);                                       // base address
    file.u64(kUndef);                                  // free-space info
    h->eof_at = file.size();
    file.u64(0);                                       // end of file (patched)
    file.u64(kUndef);                                  // driver information
    // root group symbol table entry (cache type 1: B-tree and heap addresses)
    file.u64(0);
    file.u64(root);
    file.u32(1);
    file.u32(0);
    file.u64(btree);
    file.u64(heap);
    // root object header: one symbol table message
    Bytes stab;
    stab.u64(btree);
    stab.u64(heap);
    Bytes rh = object_header({{0x11, stab}}, root, nullptr);
    file.raw(rh.b.data(), rh.size());
    // local heap: "" at 0, "Data" at 8, "Time" at 16, one free block at 24
    file.raw("HEAP", 4);
    file.u8(0);
    file.zeros(3);
    file.u64(64);
    file.u64(24);
    file.u64(heap_data);
    Bytes hd;
    hd.zeros(8);
    hd.raw("Data", 4);
    hd.zeros(4);
    hd.raw("Time", 4);
    hd.zeros(4);
    hd.u64(1);                                         // last free block
    hd.u64(40);                                        // its size
    hd.zeros(64 - hd.size());
    file.raw(hd.b.data(), hd.size());
    // B-tree (group node, leaf level, one child)
    file.raw("TREE", 4);
    file.u8(0);
    file.u8(0);
    file.u16(1);
    file.u64(kUndef);
    file.u64(kUndef);
    file.u64(0);                                       // key 0: ""
    file.u64(snod);                                    // child 0
    file.u64(16);                                      // key 1: "Time", the last name in the node
    file.zeros(btree_size - 24 - 24);
    // symbol node: "Data", "Time"
    file.raw("SNOD", 4);
    file.u8(1);
    file.u8(0);
    file.u16(2);
    const uint64_t names_off[2] = {8, 16}, ohs[2] = {data_oh, time_oh};
    for (int k = 0; k < 2; ++k) {
        file.u64(names_off[k]);
        file.u64(ohs[k]);
        file.u32(0);
        file.u32(0);
        file.zeros(16);
    }
    file.zeros(6 * 40);
    file.raw(dhdr.b.data(), dhdr.size());
    file.raw(thdr.b.data(), thdr.size());
    file.zeros(h->data_addr - file.size());
    if (std::fwrite(file.b.data(), 1, file.size(), h->f) != file.size()) {
        std::fclose(h->f);
        delete h;
        return io_fail(std::string("write failed: ") + path);
    }
    *out = h;
    return CMX_OK;
}

int cmx_h5_write_step(cmx_h5* h, double time, const double* values)
{
    if (!h || !values) return io_fail("bad argument");
    const size_t n = (size_t)h->n_nodes * h->n_vars;
    if (std::fwrite(values, sizeof(double), n, h->f) != n) return io_fail("HDF5 write failed");
    h->times.push_back(time);
    h->steps += 1;
    return CMX_OK;
}

int cmx_h5_close(cmx_h5* h)
{
    if (!h) return io_fail("bad argument");
    int rc = CMX_OK;
    const uint64_t data_bytes = h->steps * (uint64_t)h->n_nodes * h->n_vars * sizeof(double);
    const uint64_t time_addr = h->data_addr + data_bytes;
    auto put = [&](uint64_t at, uint64_t v) {
        if (std::fseek(h->f, (long)at, SEEK_SET) != 0 || std::fwrite(&v, 8, 1, h->f) != 1) rc = io_fail("HDF5 patch failed");
    };
    if (std::fseek(h->f, (long)time_addr, SEEK_SET) != 0 ||
        std::fwrite(h->times.data(), sizeof(double), h->times.size(), h->f) != h->times.size())
        rc = io_fail("HDF5 write failed");
    const uint64_t eof = time_addr + h->times.size() * sizeof(double);
    put(h->eof_at, eof);
    put(h->data_dims_at, h->steps);
    put(h->data_layout_at, h->steps ? h->data_addr : kUndef);
    put(h->data_layout_at + 8, data_bytes);
    put(h->time_dims_at, h->steps);
    put(h->time_layout_at, h->steps ? time_addr : kUndef);
    put(h->time_layout_at + 8, h->times.size() * sizeof(double));
    if (std::fclose(h->f) != 0) rc = io_fail("HDF5 close failed");
    delete h;
    return rc;
}

}  // extern "C"
