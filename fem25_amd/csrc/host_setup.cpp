// Host-side setup: quadrature/shape tables, GLL numbering, CSR pattern,
// inversion bookkeeping and element colouring.  Runs once per mesh.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "f25_internal.h"

namespace f25 {

// ---------------------------------------------------------------------------
// Gauss-Lobatto-Legendre rule on [-1, 1], ascending, computed the way
// GAUSSQ does for kind=5, alpha=beta=0, kpts=2 (gaussq.f:5-403): Golub-Welsch
// with the Lobatto modification of the last recurrence row and an implicit
// QL eigen-solve.  The same arithmetic reproduces the reference's nodes to
// the last bit (its end points are not exactly +-1), which GENID's fixed-DOF
// test (graph25.f:437-469) is sensitive to.  Contraction to FMA is off in
// this file so that the sequence of roundings is the source's.
#pragma clang fp contract(off)

namespace {
double gq_solve(double shift, int n, const double *a, const double *b) {
    double alpha = a[1] - shift;
    for (int i = 2; i <= n - 1; ++i) alpha = a[i] - shift - (b[i - 1] * b[i - 1]) / alpha;
    return 1.0 / alpha;
}

// implicit QL on the symmetric tridiagonal (d, e); z holds the first
// components of the eigenvectors (1-based arrays)
void gq_ql(int n, double *d, double *e, double *z) {
    const double machep = 2.220446049250313e-16;
    if (n == 1) return;
    e[n] = 0.0;
    for (int l = 1; l <= n; ++l) {
        int iter = 0, m;
        for (;;) {
            for (m = l; m <= n; ++m) {
                if (m == n) break;
                if (std::fabs(e[m]) <= machep * (std::fabs(d[m]) + std::fabs(d[m + 1]))) break;
            }
            double p = d[l];
            if (m == l) break;
            if (iter == 30) return;
            ++iter;
            double g = (d[l + 1] - p) / (2.0 * e[l]);
            double r = std::sqrt(g * g + 1.0);
            g = d[m] - p + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0;
            p = 0.0;
            for (int i = m - 1; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                if (std::fabs(f) < std::fabs(g)) {
                    s = f / g; r = std::sqrt(s * s + 1.0); e[i + 1] = g * r; c = 1.0 / r; s = s * c;
                } else {
                    c = g / f; r = std::sqrt(c * c + 1.0); e[i + 1] = f * r; s = 1.0 / r; c = c * s;
                }
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                f = z[i + 1];
                z[i + 1] = s * z[i] + c * f;
                z[i] = c * z[i] - s * f;
            }
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    for (int i = 1; i < n; ++i) {   // ascending, selection order as in the source
        int k = i;
        double p = d[i];
        for (int j = i + 1; j <= n; ++j)
            if (d[j] < p) { k = j; p = d[j]; }
        if (k != i) { d[k] = d[i]; d[i] = p; std::swap(z[i], z[k]); }
    }
}
}  // namespace

void gll_nodes(int n, double *x, double *w) {
    std::vector<double> t(n + 2), b(n + 2), z(n + 2);
    // Jacobi recurrence with alpha = beta = 0: diagonal zero, off-diagonal
    // sqrt(4 i^4 / ((4 i^2 - 1) 4 i^2)), zeroth moment 2
    double abi = 2.0;
    t[1] = 0.0;
    b[1] = std::sqrt(4.0 * 1.0 * 1.0 / ((abi + 1.0) * abi * abi));
    for (int i = 2; i <= n - 1; ++i) {
        abi = 2.0 * i;
        t[i] = 0.0;
        b[i] = std::sqrt(4.0 * i * (double)i * (double)i * (double)i / ((abi * abi - 1.0) * abi * abi));
    }
    t[n] = 0.0;
    const double gam = gq_solve(-1.0, n, t.data(), b.data());
    const double t1 = (-1.0 - 1.0) / (gq_solve(1.0, n, t.data(), b.data()) - gam);
    b[n - 1] = std::sqrt(t1);
    t[n] = -1.0 + gam * t1;
    z[1] = 1.0;
    gq_ql(n, t.data(), b.data(), z.data());
    for (int i = 1; i <= n; ++i) {
        x[i - 1] = t[i];
        w[i - 1] = 2.0 * z[i] * z[i];
    }
}

// Lagrange cardinal polynomial through pts with value 1 at pts[k], as
// monomial coefficients (the construction of POLCOE, dshl.f:266-314, kept
// because DSHL's rounding at high order comes from this basis).
static void cardinal_monomials(const std::vector<double> &pts, int k, std::vector<double> &cof) {
    const int n = (int)pts.size();
    std::vector<double> s(n, 0.0);
    cof.assign(n, 0.0);
    s[n - 1] = -pts[0];
    for (int i = 1; i < n; ++i) {
        for (int j = n - 1 - i; j <= n - 2; ++j) s[j] -= pts[i] * s[j + 1];
        s[n - 1] -= pts[i];
    }
    // only the j == k term of POLCOE's outer loop has a non-zero value
    double phi = (double)n;
    for (int m = n - 1; m >= 1; --m) phi = m * s[m] + pts[k] * phi;
    double ff = 1.0 / phi, b = 1.0;
    for (int m = n; m >= 1; --m) {
        cof[m - 1] += b * ff;
        b = s[m - 1] + pts[k] * b;
    }
}

static double powi(double x, int k) {
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

void make_tables(int nord, Tables &t) {
    t.nord = nord;
    t.nl = nord + 1;
    t.nen = t.nl * t.nl;
    t.nq1 = 2 * nord + 5;  // asmble.f90:110-112
    t.nq = t.nq1 * t.nq1;
    t.xipts.resize(t.nl);
    t.xq.resize(t.nq1);
    t.wq.resize(t.nq1);
    std::vector<double> wtmp(t.nl);
    gll_nodes(t.nl, t.xipts.data(), wtmp.data());
    if (t.nl % 2) t.xipts[(t.nl - 1) / 2] = 0.0;      // dshl.f:135-136
    gll_nodes(t.nq1, t.xq.data(), t.wq.data());
    if (t.nq1 % 2) t.xq[(t.nq1 - 1) / 2] = 0.0;       // dshl.f:145-146
    // values and first derivatives of the 1-D cardinal functions at the
    // quadrature points, flushed like DSHL/SHAPEF (dshl.f:169,190,259)
    const double eps_d = 2.11e-16, eps_v = 1.11e-15;
    std::vector<double> val(t.nl * t.nq1), der(t.nl * t.nq1), cof;
    for (int k = 0; k < t.nl; ++k) {
        cardinal_monomials(t.xipts, k, cof);
        for (int q = 0; q < t.nq1; ++q) {
            double x = t.xq[q], v = 0.0, d = 0.0;
            for (int i = t.nl; i >= 2; --i) v += cof[i - 1] * powi(x, i - 1);
            v += cof[0];
            if (std::fabs(v) < eps_v) v = 0.0;
            for (int i = t.nl; i >= 2; --i) d += (i - 1.0) * cof[i - 1] * powi(x, i - 2);
            if (std::fabs(d) < eps_d) d = 0.0;
            val[k * t.nq1 + q] = v;
            der[k * t.nq1 + q] = d;
        }
    }
    t.val1 = val;
    t.der1 = der;
    t.shl.assign((size_t)t.nq * t.nen * 3, 0.0);
    for (int iz = 0; iz < t.nq1; ++iz)
        for (int ix = 0; ix < t.nq1; ++ix) {
            int q = ix + t.nq1 * iz;
            for (int je = 0; je < t.nl; ++je)
                for (int jx = 0; jx < t.nl; ++jx) {
                    int a = je * t.nl + jx;      // dshl.f:149-175 node order
                    double sna = val[je * t.nq1 + iz], snb = val[jx * t.nq1 + ix];
                    double n = sna * snb, dxi = sna * der[jx * t.nq1 + ix], deta = snb * der[je * t.nq1 + iz];
                    if (std::fabs(n) < eps_d) n = 0.0;
                    if (std::fabs(dxi) < eps_d) dxi = 0.0;
                    if (std::fabs(deta) < eps_d) deta = 0.0;
                    double *o = &t.shl[((size_t)q * t.nen + a) * 3];
                    o[0] = dxi;
                    o[1] = deta;
                    o[2] = n;
                }
        }
}

// ---------------------------------------------------------------------------
// Bilinear position of GLL point (xi, eta) in element e (CSF2DND, csf2d.f:135).
void gll_xz(const f25_mesh_desc &d, int e, double xi, double eta, double &x, double &z) {
    const double eps = 2.2204e-16;
    double sf[4] = {.25 * (1.0 - xi) * (1.0 - eta), .25 * (1.0 + xi) * (1.0 - eta),
                    .25 * (1.0 + xi) * (1.0 + eta), .25 * (1.0 - xi) * (1.0 + eta)};
    x = z = 0.0;
    for (int k = 0; k < 4; ++k) {
        if (std::fabs(sf[k]) < eps) sf[k] = 0.0;
        int g = d.ieng[4 * e + k] - 1;
        x += sf[k] * d.xlocs[g];
        z += sf[k] * d.zlocs[g];
    }
}

static inline bool on_side(int ix, int ie, int nl) {
    return ix == 0 || ix == nl - 1 || ie == 0 || ie == nl - 1;
}

// GLL numbering with the semantics of GENIEN/GENID/GENLM/CNMESH.  GENIEN's
// quadratic duplicate search (graph25.f:566-575) is replaced by sorting:
// side-node occurrences are sorted by x, split into runs whose consecutive
// x differ by more than TOL (graph25.f:604), sorted by z inside each run and
// merged when x and z agree to TOL.  The resulting order is the reference's
// sorted order, so the numbers are identical.
int build_numbering(const f25_mesh_desc &d, const Tables &t, Graph &g) {
    const double tol = 1e-8;
    const int nl = t.nl, nen = t.nen, nelem = d.nelem;
    struct Occ { double x, z; int slot; };
    std::vector<Occ> occ;
    occ.reserve((size_t)nelem * 4 * (nl - 1));
    for (int e = 0; e < nelem; ++e)
        for (int ie = 0; ie < nl; ++ie)
            for (int ix = 0; ix < nl; ++ix)
                if (on_side(ix, ie, nl)) {
                    double x, z;
                    gll_xz(d, e, t.xipts[ix], t.xipts[ie], x, z);
                    occ.push_back({x, z, e * nen + ie * nl + ix});
                }
    std::stable_sort(occ.begin(), occ.end(), [](const Occ &a, const Occ &b) { return a.x < b.x; });
    g.ien.assign((size_t)nelem * nen, -5);
    // runs of x
    std::vector<size_t> run_beg;
    run_beg.push_back(0);
    for (size_t i = 1; i < occ.size(); ++i)
        if (std::fabs(occ[i].x - occ[i - 1].x) > tol) run_beg.push_back(i);
    run_beg.push_back(occ.size());
    // unique nodes per run, z-sorted
    struct Uni { double x, z; size_t first; };
    std::vector<std::vector<Uni>> runs(run_beg.size() - 1);
    std::vector<std::vector<std::pair<size_t, size_t>>> owner(run_beg.size() - 1);  // (occ index, unique idx)
    for (size_t r = 0; r + 1 < run_beg.size(); ++r) {
        std::vector<size_t> idx(run_beg[r + 1] - run_beg[r]);
        std::iota(idx.begin(), idx.end(), run_beg[r]);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
            if (occ[a].z != occ[b].z) return occ[a].z < occ[b].z;
            return occ[a].slot < occ[b].slot;
        });
        auto &u = runs[r];
        for (size_t k = 0; k < idx.size(); ++k) {
            const Occ &o = occ[idx[k]];
            if (!u.empty() && std::fabs(u.back().z - o.z) < tol && std::fabs(u.back().x - o.x) < tol) {
                owner[r].push_back({idx[k], u.size() - 1});
                if ((size_t)o.slot < u.back().first) { u.back().first = o.slot; u.back().x = o.x; u.back().z = o.z; }
            } else {
                u.push_back({o.x, o.z, (size_t)o.slot});
                owner[r].push_back({idx[k], u.size() - 1});
            }
        }
    }
    // GENIEN's z-sort loop never compares the last two x values
    // (graph25.f:592-600): a last run holding a single node is z-sorted
    // together with the run before it.
    if (runs.size() >= 2 && runs.back().size() == 1) {
        size_t r0 = runs.size() - 2;
        auto &a = runs[r0];
        const Uni last = runs.back()[0];
        size_t pos = 0;
        while (pos < a.size() && !(last.z < a[pos].z)) ++pos;
        a.insert(a.begin() + pos, last);
        for (auto &p : owner[r0])
            if (p.second >= pos) p.second++;
        for (auto &p : owner.back()) owner[r0].push_back({p.first, pos});
        runs.pop_back();
        owner.pop_back();
    }
    // number runs in order
    int next = 0;
    for (size_t r = 0; r < runs.size(); ++r) {
        std::vector<int> num(runs[r].size());
        for (size_t k = 0; k < num.size(); ++k) num[k] = ++next;
        for (auto &p : owner[r]) g.ien[occ[p.first].slot] = num[p.second];
    }
    const int nnps = next;
    int iint = nnps;
    for (int e = 0; e < nelem; ++e)
        for (int ie = 0; ie < nl; ++ie)
            for (int ix = 0; ix < nl; ++ix)
                if (!on_side(ix, ie, nl)) g.ien[(size_t)e * nen + ie * nl + ix] = ++iint;
    g.nnp = iint;
    for (int v : g.ien)
        if (v <= 0) return F25_ERR_GRAPH;

    // GENID (graph25.f:315-505): fixed DOFs on 'FL'/'FB'/'FR' sides of 'A' elements
    const int lnull = -5;
    std::vector<int> id((size_t)g.nnp * 3, lnull);
    auto fixed = [&](int g1) {
        const char *c = d.cnnpg + 2 * (g1 - 1);
        return c[0] == 'F' && (c[1] == 'L' || c[1] == 'B' || c[1] == 'R');
    };
    const double tolid = 1e-10;
    for (int e = 0; e < nelem; ++e) {
        if (d.cdomain[e] != 'A') continue;
        int icheck = 0, iav[3] = {0, 0, 0};
        for (int k = 0; k < 4 && icheck == 0; ++k) {
            int iloc = d.ieng[4 * e + k];
            if (!fixed(iloc)) continue;
            iav[icheck++] = iloc;
            for (int j = 0; j < 4 && icheck < 3; ++j) {
                int jloc = d.ieng[4 * e + j];
                if (fixed(jloc) && jloc != iloc) iav[icheck++] = jloc;
            }
        }
        if (icheck == 0) continue;
        for (int loop = 0; loop < 2; ++loop) {
            int ia1 = iav[0], ia2 = 0;
            double v1x = 0, v1z = 0;
            if (icheck > 1) {
                ia1 = iav[loop];
                ia2 = iav[loop + 1];
                if (ia2 == 0) break;
                v1x = d.xlocs[ia2 - 1] - d.xlocs[ia1 - 1];
                v1z = d.zlocs[ia2 - 1] - d.zlocs[ia1 - 1];
            } else if (loop == 1) {
                break;
            }
            for (int ie = 0; ie < nl; ++ie)
                for (int ix = 0; ix < nl; ++ix) {
                    if (!on_side(ix, ie, nl)) continue;
                    int inp = g.ien[(size_t)e * nen + ie * nl + ix] - 1;
                    double x, z;
                    gll_xz(d, e, t.xipts[ix], t.xipts[ie], x, z);
                    bool hit;
                    if (icheck == 1)
                        hit = std::fabs(x - d.xlocs[ia1 - 1]) < tolid && std::fabs(z - d.zlocs[ia1 - 1]) < tolid;
                    else
                        hit = std::fabs(v1x * (z - d.zlocs[ia1 - 1]) - v1z * (x - d.xlocs[ia1 - 1])) < tolid;
                    if (hit) id[3 * inp] = id[3 * inp + 1] = id[3 * inp + 2] = 0;
                }
        }
    }
    int idof = 0;
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nen; ++a) {
            int inp = g.ien[(size_t)e * nen + a] - 1;
            for (int i = 0; i < 3; ++i)
                if (id[3 * inp + i] == lnull) id[3 * inp + i] = ++idof;
        }
    g.ndof = idof;
    for (int v : id)
        if (v == lnull) return F25_ERR_GRAPH;
    g.lm.resize((size_t)nelem * nen * 3);
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nen; ++a)
            for (int i = 0; i < 3; ++i)
                g.lm[((size_t)e * nen + a) * 3 + i] = id[3 * (g.ien[(size_t)e * nen + a] - 1) + i];
    // CNMESH (bielak_utils.f:934-1025)
    g.cnp.assign(g.ndof, ' ');
    const int nee = 3 * nen;
    for (char dom : {'A', 'E', 'I'})
        for (int e = 0; e < nelem; ++e) {
            if (d.cdomain[e] != dom) continue;
            for (int k = 0; k < nee; ++k) {
                int dd = g.lm[(size_t)e * nee + k];
                if (dd <= 0) continue;
                char &c = g.cnp[dd - 1];
                if (dom == 'A') c = 'A';
                else if (dom == 'E') { if (c == ' ') c = 'E'; }
                else if (c == ' ') c = 'I';
                else if (c == 'E') c = 'B';
            }
        }
    return F25_OK;
}

// CSR pattern with GENCRS_QUICK's result (graph25.f:136-266): row = DOF,
// columns = sorted union of the DOFs of every element holding the row's node.
int build_csr(const f25_mesh_desc &d, const Tables &t, Graph &g) {
    const int nen = t.nen, nee = 3 * nen, nelem = d.nelem;
    // elements per fine node (CONNECT on IEN)
    std::vector<int> cnt(g.nnp + 1, 0);
    for (size_t i = 0; i < g.ien.size(); ++i) cnt[g.ien[i]]++;
    std::vector<int64_t> ptr(g.nnp + 2, 0);
    for (int n = 1; n <= g.nnp; ++n) ptr[n + 1] = ptr[n] + cnt[n];
    std::vector<int> elems(ptr[g.nnp + 1]);
    std::vector<int64_t> fill(ptr.begin(), ptr.end());
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nen; ++a) {
            int n = g.ien[(size_t)e * nen + a];
            bool dup = false;   // a node appears once per element
            for (int64_t k = ptr[n]; k < fill[n]; ++k) dup |= (elems[k] == e);
            if (!dup) elems[fill[n]++] = e;
        }
    // node of each DOF and the row pattern per node
    std::vector<int> node_of_dof(g.ndof, 0);
    for (int e = 0; e < nelem; ++e)
        for (int a = 0; a < nen; ++a)
            for (int i = 0; i < 3; ++i) {
                int dd = g.lm[((size_t)e * nen + a) * 3 + i];
                if (dd > 0) node_of_dof[dd - 1] = g.ien[(size_t)e * nen + a];
            }
    std::vector<int64_t> node_nnz(g.nnp + 1, 0);
    std::vector<std::vector<int>> pat_cache;   // not kept: patterns rebuilt in a second pass
    std::vector<int> work;
    work.reserve((size_t)nee * 8);
    auto pattern = [&](int n, std::vector<int> &w) {
        w.clear();
        for (int64_t k = ptr[n]; k < fill[n]; ++k) {
            const int *lm = &g.lm[(size_t)elems[k] * nee];
            for (int j = 0; j < nee; ++j)
                if (lm[j] > 0) w.push_back(lm[j]);
        }
        std::sort(w.begin(), w.end());
        w.erase(std::unique(w.begin(), w.end()), w.end());
    };
    for (int n = 1; n <= g.nnp; ++n) {
        pattern(n, work);
        node_nnz[n] = (int64_t)work.size();
    }
    g.irptr.assign(g.ndof + 1, 0);
    for (int r = 0; r < g.ndof; ++r) g.irptr[r + 1] = g.irptr[r] + node_nnz[node_of_dof[r]];
    g.jcptr.resize(g.irptr[g.ndof]);
    // fill rows node by node
    std::vector<std::vector<int>> dofs_of_node(1);
    std::vector<int> first_dof(g.nnp + 1, -1);
    std::vector<int> dof_next(g.ndof, -1);
    for (int r = g.ndof - 1; r >= 0; --r) {
        int n = node_of_dof[r];
        dof_next[r] = first_dof[n];
        first_dof[n] = r;
    }
    for (int n = 1; n <= g.nnp; ++n) {
        if (first_dof[n] < 0) continue;
        pattern(n, work);
        for (int r = first_dof[n]; r >= 0; r = dof_next[r])
            std::copy(work.begin(), work.end(), g.jcptr.begin() + g.irptr[r]);
    }
    g.have_csr = true;
    return F25_OK;
}

// ---------------------------------------------------------------------------
// GENGMASK (gengmask.f90:1-262, no .mask file), GRADPTRS (gengmask.f90:962-1198,
// one rank) and ELEM_WTS (gengmask.f90:758-841).
int build_inv(const f25_mesh_desc &d, const Graph &g, int nvinv, int linrec, int lkbdry,
              int nrec, const int *mrdof1, InvInfo &inv) {
    const int nnpg = d.nnpg, nelem = d.nelem;
    const int nl = d.nord + 1, nen = nl * nl;
    inv.nvinv = nvinv;
    // anchor connectivity (CONNECT on IENG): elements in increasing order
    std::vector<int> cnt(nnpg, 0);
    for (int e = 0; e < nelem; ++e)
        for (int k = 0; k < 4; ++k) cnt[d.ieng[4 * e + k] - 1]++;
    inv.ncon = *std::max_element(cnt.begin(), cnt.end());
    inv.mconn.assign((size_t)inv.ncon * nnpg, 0);
    std::vector<int> fill(nnpg, 0);
    for (int e = 0; e < nelem; ++e)
        for (int k = 0; k < 4; ++k) {
            int a = d.ieng[4 * e + k] - 1;
            inv.mconn[(size_t)fill[a]++ * nnpg + a] = e + 1;
        }
    auto is = [&](int a, const char *s) { return d.cnnpg[2 * a] == s[0] && d.cnnpg[2 * a + 1] == s[1]; };
    inv.maskg.assign(nnpg, 0);
    int n = 0;
    bool shift = false;
    for (int a = 0; a < nnpg; ++a) {
        int e = inv.mconn[a];
        bool keep = !(e > 0 && d.cdomain[e - 1] != 'I');
        if (keep && (is(a, "II") || is(a, "BI") || is(a, "FS"))) inv.maskg[a] = ++n;
    }
    inv.nnpinv = n;
    const int corner_node[4] = {0, nl - 1, nl * nl - 1, nl * (nl - 1)};   // IA2IAE, vtk25.f:643
    for (int e = 0; e < nelem; ++e)
        for (int k = 0; k < 4; ++k) {
            int a = d.ieng[4 * e + k] - 1;
            if (inv.maskg[a] <= 0) continue;
            const int *lm = &g.lm[((size_t)e * nen + corner_node[k]) * 3];
            bool drop = lm[0] <= 0 || lm[1] <= 0 || lm[2] <= 0;
            if (!drop && !linrec)
                for (int r = 0; r < nrec && !drop; ++r) drop = (lm[0] == mrdof1[r]);
            if (drop) { inv.maskg[a] = 0; inv.nnpinv--; shift = true; }
        }
    if (!lkbdry)
        for (int e = 0; e < nelem; ++e) {
            bool bdr = false;
            for (int k = 0; k < 4; ++k) {
                int a = d.ieng[4 * e + k] - 1;
                if (!is(a, "II") && !is(a, "FS")) bdr = true;
            }
            if (bdr)
                for (int k = 0; k < 4; ++k) {
                    int a = d.ieng[4 * e + k] - 1;
                    if (inv.maskg[a] > 0) { inv.maskg[a] = 0; inv.nnpinv--; shift = true; }
                }
        }
    if (inv.nnpinv < 1) return F25_ERR_ARG;
    if (shift) {
        n = 0;
        for (int a = 0; a < nnpg; ++a)
            if (inv.maskg[a] > 0) inv.maskg[a] = ++n;
    }
    inv.na35 = inv.nnpinv * nvinv;
    // GRADPTRS
    inv.nnpgl = inv.nnpinv * nvinv;
    inv.jcsc.assign(inv.nnpgl + 1, 1);
    inv.mygrad.assign(inv.nnpgl, 0);
    inv.node_of_col.assign(inv.nnpgl, 0);
    inv.icsc.clear();
    inv.mbufrhs = 0;
    std::vector<int> rows;
    int col = 0;
    const int nee = 3 * nen;
    for (int a = 0; a < nnpg; ++a) {
        int ip = inv.maskg[a];
        if (ip <= 0) continue;
        rows.clear();
        for (int c = 0; c < inv.ncon; ++c) {
            int e = inv.mconn[(size_t)c * nnpg + a];
            if (e <= 0) break;
            for (int j = 0; j < nee; ++j) {
                int dd = g.lm[(size_t)(e - 1) * nee + j];
                if (dd > 0) rows.push_back(dd);
            }
        }
        std::sort(rows.begin(), rows.end());
        rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
        inv.mbufrhs = std::max(inv.mbufrhs, (int)rows.size());
        for (int v = 1; v <= nvinv; ++v) {
            inv.icsc.insert(inv.icsc.end(), rows.begin(), rows.end());
            inv.mygrad[col] = (ip - 1) * nvinv + v;
            inv.node_of_col[col] = a;
            inv.jcsc[col + 1] = (int)inv.icsc.size() + 1;
            ++col;
        }
    }
    inv.nz_fdist = (int)inv.icsc.size();
    // ELEM_WTS
    inv.elem_wts.assign(nelem, 0.0);
    std::vector<char> init(nelem, 0);
    double wtmax = 0.0;
    for (int a = 0; a < nnpg; ++a) {
        if (inv.maskg[a] <= 0) continue;
        for (int c = 0; c < inv.ncon; ++c) {
            int e = inv.mconn[(size_t)c * nnpg + a];
            if (e <= 0) break;
            --e;
            if (init[e]) continue;
            double r = 0.0;
            for (int k = 0; k < 4; ++k) {
                int j1 = d.ieng[4 * e + k] - 1, j2 = d.ieng[4 * e + (k + 1) % 4] - 1;
                r += d.xlocs[j1] * d.zlocs[j2] - d.xlocs[j2] * d.zlocs[j1];
            }
            r = 0.5 * std::fabs(r);
            inv.elem_wts[e] = r;
            init[e] = 1;
            wtmax = std::max(wtmax, r);
        }
    }
    for (int e = 0; e < nelem; ++e)
        if (inv.elem_wts[e] > 0.0) inv.elem_wts[e] = wtmax / inv.elem_wts[e];
    return wtmax > 0.0 ? F25_OK : F25_ERR_ARG;
}

// Greedy colouring in element order over the anchor-sharing graph.
int colour_elements(int nelem, int nnpg, const int *ieng, std::vector<int> &order,
                    std::vector<int> &colour_ptr) {
    std::vector<std::vector<int>> at(nnpg);
    for (int e = 0; e < nelem; ++e)
        for (int k = 0; k < 4; ++k) at[ieng[4 * e + k] - 1].push_back(e);
    std::vector<int> col(nelem, -1);
    int ncol = 0;
    for (int e = 0; e < nelem; ++e) {
        uint64_t used = 0;
        for (int k = 0; k < 4; ++k)
            for (int f : at[ieng[4 * e + k] - 1])
                if (col[f] >= 0 && col[f] < 64) used |= (1ull << col[f]);
        int c = 0;
        while (used & (1ull << c)) ++c;
        col[e] = c;
        ncol = std::max(ncol, c + 1);
    }
    colour_ptr.assign(ncol + 1, 0);
    for (int e = 0; e < nelem; ++e) colour_ptr[col[e] + 1]++;
    for (int c = 0; c < ncol; ++c) colour_ptr[c + 1] += colour_ptr[c];
    order.resize(nelem);
    std::vector<int> pos(colour_ptr.begin(), colour_ptr.end() - 1);
    for (int e = 0; e < nelem; ++e) order[pos[col[e]]++] = e;
    return ncol;
}

}  // namespace f25
