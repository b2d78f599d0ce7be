// Host-side setup of the monodomain operators: slab meshes, P1 assembly and
// the sparse formats the device kernels read.  Runs once per context (the
// reference assembles once too: MonodomainSolver has mMatrixIsConstant,
// heart/src/solver/electrics/monodomain/MonodomainSolver.cpp:190).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

#include "cmx_internal.h"

namespace cmx {

namespace {

// Inverse-transpose gradients of the P1 basis on one simplex and its measure.
// grads[a][r] = d phi_a / d x_r.  Returns false for a degenerate element.
bool simplex_geometry(int dim, const double* X /*[(dim+1)][dim]*/, double grads[4][3], double& measure)
{
    double J[3][3] = {};
    for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) J[r][c] = X[(c + 1) * dim + r] - X[r];
    double inv[3][3] = {};
    double det;
    if (dim == 1) {
        det = J[0][0];
        if (det == 0.0) return false;
        inv[0][0] = 1.0 / det;
    } else if (dim == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        if (det == 0.0) return false;
        inv[0][0] = J[1][1] / det;  inv[0][1] = -J[0][1] / det;
        inv[1][0] = -J[1][0] / det; inv[1][1] = J[0][0] / det;
    } else {
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        if (det == 0.0) return false;
        inv[0][0] = c00 / det;
        inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        inv[1][0] = c01 / det;
        inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        inv[2][0] = c02 / det;
        inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    }
    // reference gradients: phi_0 = 1 - sum xi, phi_k = xi_k
    for (int r = 0; r < dim; ++r) {
        double s = 0;
        for (int c = 0; c < dim; ++c) s += inv[c][r];
        grads[0][r] = -s;
        for (int a = 1; a <= dim; ++a) grads[a][r] = inv[a - 1][r];
    }
    double fact = 1.0;
    for (int d = 2; d <= dim; ++d) fact *= d;
    measure = std::fabs(det) / fact;
    return true;
}

}  // namespace

bool assemble_p1(const cmx_mesh& mesh, bool lumped, HostCsr& out, std::string& err)
{
    const int dim = mesh.dim, n = mesh.n_nodes, ne = mesh.n_elems;
    // elements of the node dimension, or line elements (cables) in 2-D / 3-D
    const int edim = mesh.elem_dim > 0 ? mesh.elem_dim : dim, nb = edim + 1;
    if (dim < 1 || dim > 3 || n <= 0 || ne <= 0 || !mesh.coords || !mesh.elems) {
        err = "mesh: bad dimension, sizes or null arrays";
        return false;
    }
    if (edim != dim && edim != 1) {
        err = "mesh: elements must have the node dimension or be lines";
        return false;
    }
    if (edim != dim && mesh.sigma_elem) {
        err = "mesh: line elements take the scalar conductivity sigma[0], not per-element tensors";
        return false;
    }
    for (long e = 0; e < (long)ne * nb; ++e)
        if (mesh.elems[e] < 0 || mesh.elems[e] >= n) {
            err = "mesh: element references node " + std::to_string(mesh.elems[e]) + " out of range";
            return false;
        }
    // Sparsity: nodes sharing an element.
    std::vector<int32_t> deg(n, 0);
    std::vector<std::vector<int32_t>> adj(n);
    for (int e = 0; e < ne; ++e) {
        const int32_t* nd = mesh.elems + (size_t)e * nb;
        for (int a = 0; a < nb; ++a)
            for (int b = 0; b < nb; ++b) adj[nd[a]].push_back(nd[b]);
    }
    out.n = n;
    out.row_ptr.assign(n + 1, 0);
    out.max_row = 0;
    for (int i = 0; i < n; ++i) {
        auto& r = adj[i];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        out.row_ptr[i + 1] = out.row_ptr[i] + (int32_t)r.size();
        out.max_row = std::max<int>(out.max_row, (int)r.size());
    }
    const size_t nnz = out.row_ptr[n];
    out.col.resize(nnz);
    for (int i = 0; i < n; ++i) std::copy(adj[i].begin(), adj[i].end(), out.col.begin() + out.row_ptr[i]);
    adj.clear();
    adj.shrink_to_fit();
    out.mass.assign(nnz, 0.0);
    out.stiff.assign(nnz, 0.0);

    const double mass_scale = 1.0 / ((edim + 1) * (edim + 2));
    // a line's conductivity acts along it: sigma[0] I gives g_a . sigma g_b = sigma[0] g_a . g_b
    double line_sigma[9] = {};
    for (int r = 0; r < dim; ++r) line_sigma[r * dim + r] = mesh.sigma[0];
    for (int e = 0; e < ne; ++e) {
        const int32_t* nd = mesh.elems + (size_t)e * nb;
        double X[4 * 3];
        for (int a = 0; a < nb; ++a)
            for (int r = 0; r < dim; ++r) X[a * dim + r] = mesh.coords[(size_t)nd[a] * dim + r];
        double g[4][3] = {}, vol;
        if (edim == dim) {
            if (!simplex_geometry(dim, X, g, vol)) {
                err = "mesh: element " + std::to_string(e) + " is degenerate";
                return false;
            }
        } else {
            // line element: length L, basis gradients -t/L and t/L along its unit tangent t
            double t[3] = {}, L2 = 0;
            for (int r = 0; r < dim; ++r) {
                t[r] = X[dim + r] - X[r];
                L2 += t[r] * t[r];
            }
            vol = std::sqrt(L2);
            if (!(vol > 0)) {
                err = "mesh: element " + std::to_string(e) + " is degenerate";
                return false;
            }
            for (int r = 0; r < dim; ++r) {
                g[0][r] = -t[r] / L2;
                g[1][r] = t[r] / L2;
            }
        }
        const double* sig = edim != dim ? line_sigma : mesh.sigma_elem ? mesh.sigma_elem + (size_t)e * dim * dim : mesh.sigma;
        // (sigma G^T) columns
        double sg[4][3];
        for (int b = 0; b < nb; ++b)
            for (int r = 0; r < dim; ++r) {
                double s = 0;
                for (int c = 0; c < dim; ++c) s += sig[r * dim + c] * g[b][c];
                sg[b][r] = s;
            }
        for (int a = 0; a < nb; ++a) {
            const int i = nd[a];
            const int32_t* cb = out.col.data() + out.row_ptr[i];
            const int32_t* ce = out.col.data() + out.row_ptr[i + 1];
            for (int b = 0; b < nb; ++b) {
                const size_t p = std::lower_bound(cb, ce, nd[b]) - out.col.data();
                double kab = 0;
                for (int r = 0; r < dim; ++r) kab += g[a][r] * sg[b][r];
                out.stiff[p] += vol * kab;
                if (lumped) {
                    // row sum of the consistent element mass, onto the diagonal
                    // (pde/src/solver/MassMatrixAssembler.hpp:88-101)
                    if (a == b) out.mass[p] += vol * mass_scale * (nb + 1);
                } else {
                    out.mass[p] += vol * mass_scale * (a == b ? 2.0 : 1.0);
                }
            }
        }
    }
    return true;
}

bool build_stencil(const HostCsr& A, int max_patterns, int max_width, StencilTable& out)
{
    const int n = A.n;
    std::vector<int32_t> uni;
    for (int i = 0; i < n; ++i)
        for (int q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q) uni.push_back(A.col[q] - i);
    std::sort(uni.begin(), uni.end());
    uni.erase(std::unique(uni.begin(), uni.end()), uni.end());
    if ((int)uni.size() > max_width) return false;
    const int W = (int)uni.size();
    out.width = W;
    out.off = uni;
    out.pid.assign(n, 0);
    out.m.clear();
    out.k.clear();
    // candidates bucketed by the set of offsets a row actually has
    std::map<std::vector<int>, std::vector<int>> by_shape;
    std::vector<double> rm(W), rk(W);
    std::vector<int> shape;
    int count = 0;
    for (int i = 0; i < n; ++i) {
        std::fill(rm.begin(), rm.end(), 0.0);
        std::fill(rk.begin(), rk.end(), 0.0);
        shape.clear();
        double scale_m = 0, scale_k = 0;
        for (int q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q) {
            const int s = (int)(std::lower_bound(uni.begin(), uni.end(), A.col[q] - i) - uni.begin());
            rm[s] = A.mass[q];
            rk[s] = A.stiff[q];
            shape.push_back(s);
            scale_m = std::max(scale_m, std::fabs(A.mass[q]));
            scale_k = std::max(scale_k, std::fabs(A.stiff[q]));
        }
        const double tm = 1e-12 * scale_m, tk = 1e-12 * scale_k;
        auto& cands = by_shape[shape];
        int found = -1;
        for (int c : cands) {
            bool same = true;
            for (int s = 0; s < W && same; ++s) {
                const size_t d = (size_t)c * W + s;
                same = std::fabs(out.m[d] - rm[s]) <= tm && std::fabs(out.k[d] - rk[s]) <= tk;
            }
            if (same) { found = c; break; }
        }
        if (found < 0) {
            if (count >= max_patterns) return false;
            found = count++;
            cands.push_back(found);
            out.m.insert(out.m.end(), rm.begin(), rm.end());
            out.k.insert(out.k.end(), rk.begin(), rk.end());
        }
        out.pid[i] = (uint8_t)found;
    }
    out.n_patterns = count;
    return true;
}

void build_ell(const HostCsr& A, EllMatrix& out)
{
    const int n = A.n, W = A.max_row;
    out.width = W;
    out.col.assign(ell_size(n, W), 0);
    out.m.assign(ell_size(n, W), 0.0);
    out.k.assign(ell_size(n, W), 0.0);
    for (int i = 0; i < n; ++i) {
        const int b = A.row_ptr[i], e = A.row_ptr[i + 1];
        for (int s = 0; s < W; ++s) {
            const size_t d = ell_index(i, s, W);
            if (b + s < e) {
                out.col[d] = A.col[b + s];
                out.m[d] = A.mass[b + s];
                out.k[d] = A.stiff[b + s];
            } else {
                out.col[d] = i;   // padding: self, zero value
            }
        }
    }
}

bool build_offset_ell(const HostCsr& A, int max_sets, OffsetEll& out)
{
    constexpr size_t kMaxOffsetLists = 1 << 16;
    const int n = A.n, W = A.max_row;
    // distinct offset lists, and the rows that have each
    std::map<std::vector<int32_t>, int> list_id;
    std::vector<std::vector<int32_t>> lists;
    std::vector<int> row_list(n);
    std::vector<int32_t> offs;
    for (int i = 0; i < n; ++i) {
        offs.clear();
        for (int p = A.row_ptr[i]; p < A.row_ptr[i + 1]; ++p) offs.push_back(A.col[p] - i);
        std::sort(offs.begin(), offs.end());
        auto it = list_id.find(offs);
        if (it == list_id.end()) {
            // a mesh with this many distinct rows will not fit the set table
            if (lists.size() >= kMaxOffsetLists) return false;
            it = list_id.emplace(offs, (int)lists.size()).first;
            lists.push_back(offs);
        }
        row_list[i] = it->second;
    }
    // sets: longest lists first, each list into the first set containing it
    std::vector<int> order(lists.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lists[a].size() > lists[b].size(); });
    std::vector<std::vector<int32_t>> sets;
    std::vector<int> list_set(lists.size(), -1);
    for (int l : order) {
        for (size_t q = 0; q < sets.size() && list_set[l] < 0; ++q)
            if (std::includes(sets[q].begin(), sets[q].end(), lists[l].begin(), lists[l].end())) list_set[l] = (int)q;
        if (list_set[l] < 0) {
            if ((int)sets.size() >= max_sets) return false;
            list_set[l] = (int)sets.size();
            sets.push_back(lists[l]);
        }
    }
    out.width = W;
    out.n_sets = (int)sets.size();
    out.offsets.assign((size_t)out.n_sets * W, 0);
    out.max_abs = 0;
    for (int q = 0; q < out.n_sets; ++q)
        for (size_t k = 0; k < sets[q].size(); ++k) {
            out.offsets[(size_t)q * W + k] = sets[q][k];
            out.max_abs = std::max(out.max_abs, std::abs(sets[q][k]));
        }
    out.set.assign(n, 0);
    out.m.assign(ell_size(n, W), 0.0);
    out.k.assign(ell_size(n, W), 0.0);
    for (int i = 0; i < n; ++i) {
        const int q = list_set[row_list[i]];
        out.set[i] = (uint8_t)q;
        const std::vector<int32_t>& so = sets[q];
        for (int p = A.row_ptr[i]; p < A.row_ptr[i + 1]; ++p) {
            const int32_t o = A.col[p] - i;
            const int k = (int)(std::lower_bound(so.begin(), so.end(), o) - so.begin());
            out.m[ell_index(i, k, W)] = A.mass[p];
            out.k[ell_index(i, k, W)] = A.stiff[p];
        }
    }
    return true;
}

int slab_mesh(int dim, int ex, int ey, int ez, double h, int* n_nodes, int* n_elems,
              double* coords, int32_t* elems)
{
    if (h <= 0 || ex <= 0 || (dim >= 2 && ey <= 0) || (dim >= 3 && ez <= 0)) return CMX_E_ARG;
    long nn, ne;
    if (dim == 1) { nn = ex + 1; ne = ex; }
    else if (dim == 2) { nn = (long)(ex + 1) * (ey + 1); ne = 2L * ex * ey; }
    else if (dim == 3) { nn = (long)(ex + 1) * (ey + 1) * (ez + 1); ne = 6L * ex * ey * ez; }
    else return CMX_E_ARG;
    if (nn > INT32_MAX || ne > INT32_MAX) return CMX_E_ARG;
    if (n_nodes) *n_nodes = (int)nn;
    if (n_elems) *n_elems = (int)ne;
    if (!coords || !elems) return CMX_OK;
    if (dim == 1) {
        for (int i = 0; i <= ex; ++i) coords[i] = i * h;
        for (int e = 0; e < ex; ++e) { elems[2 * e] = e; elems[2 * e + 1] = e + 1; }
    } else if (dim == 2) {
        // Staggered: the diagonal of square (i, j) follows the parity of
        // i + (ey - j) (AbstractTetrahedralMesh.cpp:349-376).
        for (int j = 0; j <= ey; ++j)
            for (int i = 0; i <= ex; ++i) {
                const long q = (long)j * (ex + 1) + i;
                coords[2 * q] = i * h;
                coords[2 * q + 1] = j * h;
            }
        long e = 0;
        for (int j = 0; j < ey; ++j)
            for (int i = 0; i < ex; ++i) {
                const bool odd = ((i + (ey - j)) % 2) == 1;
                const int32_t nw = (j + 1) * (ex + 1) + i, sw = j * (ex + 1) + i;
                int32_t* t = elems + 3 * e++;
                t[0] = nw; t[1] = nw + 1; t[2] = odd ? sw + 1 : sw;
                t = elems + 3 * e++;
                t[0] = sw + 1; t[1] = sw; t[2] = odd ? nw : nw + 1;
            }
    } else {
        // Six tetrahedra around the cube diagonal from local corner 0 to 7,
        // corner c = x + 2y + 4z (AbstractTetrahedralMesh.cpp:404-440).
        static const int kTets[6][4] = {{0, 1, 5, 7}, {0, 1, 3, 7}, {0, 2, 3, 7},
                                        {0, 2, 6, 7}, {0, 4, 6, 7}, {0, 4, 5, 7}};
        for (int k = 0; k <= ez; ++k)
            for (int j = 0; j <= ey; ++j)
                for (int i = 0; i <= ex; ++i) {
                    const long q = ((long)k * (ey + 1) + j) * (ex + 1) + i;
                    coords[3 * q] = i * h;
                    coords[3 * q + 1] = j * h;
                    coords[3 * q + 2] = k * h;
                }
        long e = 0;
        for (int k = 0; k < ez; ++k)
            for (int j = 0; j < ey; ++j)
                for (int i = 0; i < ex; ++i) {
                    int32_t corner[8];
                    for (int c = 0; c < 8; ++c)
                        corner[c] = (i + (c & 1)) + (ex + 1) * ((j + ((c >> 1) & 1)) + (ey + 1) * (k + (c >> 2)));
                    for (int t = 0; t < 6; ++t, ++e)
                        for (int v = 0; v < 4; ++v) elems[4 * e + v] = corner[kTets[t][v]];
                }
    }
    return CMX_OK;
}

}  // namespace cmx
