// Internal declarations shared by the host setup code and the HIP kernels.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "fem25hip.h"

namespace f25 {

// Quadrature and interpolation tables for one polynomial order (DSHL
// semantics, dshl.f:1-243): nl = p+1 GLL interpolation points, nq1 = 2p+5
// GLL quadrature points per direction (asmble.f90:110-112).
struct Tables {
    int nord = 0, nl = 0, nen = 0, nq1 = 0, nq = 0;
    std::vector<double> xipts;   // nl
    std::vector<double> xq, wq;  // nq1
    // shl[(q*nen + a)*3 + c], q = ix + nq1*iz, c: 0 dN/dxi, 1 dN/deta, 2 N
    std::vector<double> shl;
    // the 1-D cardinal functions and derivatives at the quadrature points,
    // val1[k*nq1 + q], der1[k*nq1 + q] (shl is their tensor product)
    std::vector<double> val1, der1;
};
void make_tables(int nord, Tables &t);
void gll_nodes(int n, double *x, double *w);

// Host-side mesh graph (GEN_GRAPH25 numbering, graph25_driver.f90:30-86).
struct Graph {
    int nnp = 0, ndof = 0, ncon = 0;
    std::vector<int> ien;        // (nelem, nen) 1-based fine-node numbers
    std::vector<int> lm;         // (nelem, nen, 3) 1-based DOFs, 0 fixed
    std::vector<int64_t> irptr;  // ndof+1, 0-based offsets
    std::vector<int> jcptr;      // nz, 1-based column DOFs (sorted per row)
    std::vector<char> cnp;       // ndof
    bool have_csr = false;
};
int build_numbering(const f25_mesh_desc &d, const Tables &t, Graph &g);   // IEN, ID, LM, CNP
// position of GLL point (xi, eta) of element e by CSF2DND (csf2d.f:135)
void gll_xz(const f25_mesh_desc &d, int e, double xi, double eta, double &x, double &z);
int build_csr(const f25_mesh_desc &d, const Tables &t, Graph &g);         // IRPTR/JCPTR

// Inversion bookkeeping (GENGMASK, GRADPTRS, ELEM_WTS, one rank).
struct InvInfo {
    int nvinv = 1, nnpinv = 0, na35 = 0, ncon = 0, nnpgl = 0, nz_fdist = 0, mbufrhs = 0;
    bool lisp = true, liss = false;   // PP: P; SS: S; PS: both (NVINV = 2)
    double vpvs = 1.7320508075688772;
    std::vector<int> mconn;      // (ncon, nnpg): mconn[c*nnpg + g], 1-based elements, 0 none
    std::vector<int> maskg;      // nnpg
    std::vector<int> mygrad;     // nnpgl
    std::vector<int> jcsc;       // nnpgl+1, 1-based
    std::vector<int> icsc;       // nz_fdist, 1-based DOFs
    std::vector<double> elem_wts;// nelem
    std::vector<int> node_of_col;// nnpgl: anchor (0-based) of each column
    // second-order products (QPMFREE): elements with at least one local
    // column on a corner, in colour order, and per element the column on
    // each corner (-1 none); built on first use
    std::vector<int> dab_elem, dab_cptr, dab_col4;
};
int build_inv(const f25_mesh_desc &d, const Graph &g, int nvinv, int linrec, int lkbdry,
              int nrec, const int *mrdof1, InvInfo &inv);

// Greedy element colouring: elements of one colour share no anchor (and
// therefore no GLL node) on a conforming quadrilateral mesh.
int colour_elements(int nelem, int nnpg, const int *ieng, std::vector<int> &order,
                    std::vector<int> &colour_ptr);

}  // namespace f25
