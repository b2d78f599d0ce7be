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

// ELL: width slots per row, column-major ([slot][n]) for coalesced loads.
struct EllMatrix {
    int width = 0;
    std::vector<int32_t> col;      // [width][n]
    std::vector<double> m, k;      // [width][n]
};
void build_ell(const HostCsr& A, EllMatrix& out);

int slab_mesh(int dim, int ex, int ey, int ez, double h, int* n_nodes, int* n_elems,
              double* coords, int32_t* elems);

}  // namespace cmx
