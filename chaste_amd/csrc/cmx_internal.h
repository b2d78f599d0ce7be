// Internal declarations shared by the host assembly (cmx_host.cpp) and the
// device runtime (cmx_device.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "cmx.h"

namespace cmx {

// Assembled P1 operators in CSR, columns sorted per row.
//   M = mass matrix (unscaled), K = stiffness (div sigma grad).
struct HostCsr {
    int n = 0;
    int max_row = 0;
    std::vector<int32_t> row_ptr, col;
    std::vector<double> mass, stiff;
};

// Element-by-element P1 assembly with closed-form element integrals:
//   M_e[a][b] = |T| (1 + delta_ab) / ((d+1)(d+2)),  K_e = |T| G sigma G^T,
// the values the reference's 2nd-order Gauss rule produces exactly
// (pde/src/solver/AbstractFeVolumeIntegralAssembler.hpp:345-459,
//  pde/src/solver/MassMatrixAssembler.hpp:76-118,
//  heart/src/solver/electrics/monodomain/MonodomainStiffnessMatrixAssembler.hpp:68-84).
bool assemble_p1(const cmx_mesh& mesh, bool lumped, HostCsr& out, std::string& err);

// Uniform-offset stencil: every row reads the same column offsets (the union
// of all rows' offsets, i.e. the mesh stencil) and only the values depend on
// the row.  Rows whose value vectors over that union agree to a relative
// 1e-12 of the row scale share one pattern.  On a regular slab there are at
// most 27 (3-D) or a few dozen (2-D) patterns, so the SpMV reads an 8-bit
// pattern id per row, the offsets come from kernel arguments and the values
// from a table staged in LDS.  Entries absent from a row hold value 0.
struct StencilTable {
    int width = 0;                 // |offset union|
    int n_patterns = 0;
    std::vector<int32_t> off;      // [width], sorted
    std::vector<uint8_t> pid;      // [n]
    std::vector<double> m, k;      // [n_patterns][width]
};
bool build_stencil(const HostCsr& A, int max_patterns, int max_width, StencilTable& out);

// ELL tables are sliced: rows in slices of 64 (one wavefront), each slice
// holding its width slots one after another, 64 values each ([n/64][width][64],
// the last slice padded).  A wavefront reads slot k of its 64 rows as one
// coalesced 256-byte (FP32) line run, and its whole row block from one
// contiguous span instead of one span per slot plane.  EllRows::vidx on the
// device computes the same index.
constexpr int kEllSlice = 64;
inline size_t ell_index(int i, int k, int width)
{
    return ((size_t)(i / kEllSlice) * width + k) * kEllSlice + (size_t)(i % kEllSlice);
}
inline size_t ell_size(int n, int width)
{
    return (size_t)((n + kEllSlice - 1) / kEllSlice) * width * kEllSlice;
}

// ELL: width slots per row, sliced as above.
struct EllMatrix {
    int width = 0;
    std::vector<int32_t> col;      // ell_size(n, width), ell_index order
    std::vector<double> m, k;      // ell_size(n, width), ell_index order
};
void build_ell(const HostCsr& A, EllMatrix& out);

// ELL with per-row offset sets: row i's slot k holds the entry of column
// i + offsets[set[i]][k] (zero where the row has no such entry), so a row
// carries one set id instead of a column index per slot.  Each row is
// assigned a set holding all its offsets; sets come from the rows' distinct
// offset lists, longest first, a list joining an earlier set that contains it.
// Returns false (out untouched) when more than max_sets sets are needed.
struct OffsetEll {
    int width = 0, n_sets = 0;
    int max_abs = 0;                    // largest |offset| of any set
    std::vector<int32_t> offsets;       // [n_sets][width] (sets shorter than width repeat offset 0 as padding)
    std::vector<uint8_t> set;           // [n]
    std::vector<double> m, k;           // ell_size(n, width), ell_index order
};
bool build_offset_ell(const HostCsr& A, int max_sets, OffsetEll& out);

int slab_mesh(int dim, int ex, int ey, int ez, double h, int* n_nodes, int* n_elems,
              double* coords, int32_t* elems);

}  // namespace cmx
