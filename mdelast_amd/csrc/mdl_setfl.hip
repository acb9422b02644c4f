// Native SETFL loader and quintic-spline node derivatives (host-only code).
//
// Reading follows get_EAM_alloy_file (get_potential.py:58-147 of mdelast v1.2);
// the node derivatives are those of get_derivatives (get_potential.py:49-55),
// which asks SciPy for the k = 5 interpolating spline, s = 0. That spline is
// unique once its knots are fixed: FITPACK places the interior knots at the data
// points x[3..n-4] (the "not-a-knot" choice for odd degree), with six-fold end
// knots. It is rebuilt here from the B-spline collocation equations:
//
//   sum_j c_j B_j(x_i) = y_i,  i = 0..n-1,
//
// a banded, totally positive system (de Boor, A Practical Guide to Splines,
// ch. XIII), so Gaussian elimination without pivoting is stable. The basis
// values come from the Cox-de Boor recurrence (Piegl & Tiller, The NURBS Book,
// algorithm A2.2), the derivatives from the differenced coefficients
// c'_j = p (c_{j+1} - c_j) / (t_{j+p+1} - t_{j+1}) on the knot vector without
// its end knots. Solve and differencing run in long double.

#include "mdl_internal.hpp"

#include "../../include/mdelast_setfl.h"

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace mdl {
namespace {

using ld = long double;

// units.hpp:6-24 (the driver's own constants)
constexpr double EV = 1.6021773e-12;
constexpr double ANGSTROM = 1e-8;

// Knot span l of x in t (t[l] <= x < t[l+1]) for a spline of degree p with
// ncoef coefficients; the right end belongs to the last span.
int64_t find_span(const std::vector<ld> &t, int p, int64_t ncoef, ld x) {
    if (x >= t[ncoef])
        return ncoef - 1;
    int64_t lo = p, hi = ncoef;   // t[lo] <= x < t[hi]
    while (hi - lo > 1) {
        int64_t const mid = (lo + hi) / 2;
        if (x < t[mid])
            hi = mid;
        else
            lo = mid;
    }
    return lo;
}

// Non-zero basis functions N[0..p] of degree p at x in span l (A2.2).
void basis(const std::vector<ld> &t, int p, int64_t l, ld x, ld *N) {
    ld left[6], right[6];
    N[0] = 1.L;
    for (int j = 1; j <= p; ++j) {
        left[j] = x - t[l + 1 - j];
        right[j] = t[l + j] - x;
        ld saved = 0.L;
        for (int r = 0; r < j; ++r) {
            ld const tmp = N[r] / (right[r + 1] + left[j - r]);
            N[r] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        N[j] = saved;
    }
}

ld spline_eval(const std::vector<ld> &t, int p, const std::vector<ld> &c, ld x) {
    int64_t const ncoef = (int64_t)c.size();
    int64_t const l = find_span(t, p, ncoef, x);
    ld N[6];
    basis(t, p, l, x, N);
    ld s = 0.L;
    for (int r = 0; r <= p; ++r)
        s += c[l - p + r] * N[r];
    return s;
}

// Coefficients and knots of the derivative of a degree-p spline.
void differentiate(const std::vector<ld> &t, int p, const std::vector<ld> &c,
                   std::vector<ld> &td, std::vector<ld> &cd) {
    int64_t const n = (int64_t)c.size();
    cd.resize(n - 1);
    for (int64_t j = 0; j + 1 < n; ++j)
        cd[j] = p * (c[j + 1] - c[j]) / (t[j + p + 1] - t[j + 1]);
    td.assign(t.begin() + 1, t.end() - 1);
}

bool spline5_derivatives(int64_t n, const double *x, const double *y, double *d1, double *d2) {
    constexpr int P = 5;
    // knots: x0 (6 times), x[3..n-4], x[n-1] (6 times)
    std::vector<ld> t;
    t.reserve(n + P + 1);
    for (int k = 0; k <= P; ++k)
        t.push_back(x[0]);
    for (int64_t i = 3; i <= n - 4; ++i)
        t.push_back(x[i]);
    for (int k = 0; k <= P; ++k)
        t.push_back(x[n - 1]);
    // banded collocation matrix: row i holds columns l_i - 5 .. l_i; stored as
    // band[i][col - i + KL] with KL = KU = 5
    constexpr int KL = P, KU = P, W = KL + KU + 1;
    std::vector<ld> band((size_t)n * W, 0.L), rhs(n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t const l = find_span(t, P, n, x[i]);
        ld N[6];
        basis(t, P, l, x[i], N);
        for (int r = 0; r <= P; ++r) {
            int64_t const col = l - P + r;
            int64_t const off = col - i + KL;
            if (off < 0 || off >= W)
                return false;   // cannot happen for strictly increasing x
            band[(size_t)i * W + off] = N[r];
        }
        rhs[i] = y[i];
    }
    auto A = [&](int64_t i, int64_t j) -> ld & { return band[(size_t)i * W + (j - i + KL)]; };
    // Gaussian elimination without pivoting (totally positive matrix)
    for (int64_t k = 0; k < n; ++k) {
        ld const piv = A(k, k);
        if (piv == 0.L)
            return false;
        int64_t const iend = std::min<int64_t>(n - 1, k + KL);
        int64_t const jend = std::min<int64_t>(n - 1, k + KU);
        for (int64_t i = k + 1; i <= iend; ++i) {
            ld const f = A(i, k) / piv;
            if (f == 0.L)
                continue;
            for (int64_t j = k; j <= jend; ++j)
                if (j - i + KL < W)
                    A(i, j) -= f * A(k, j);
            rhs[i] -= f * rhs[k];
        }
    }
    std::vector<ld> c(n);
    for (int64_t i = n - 1; i >= 0; --i) {
        ld s = rhs[i];
        int64_t const jend = std::min<int64_t>(n - 1, i + KU);
        for (int64_t j = i + 1; j <= jend; ++j)
            s -= A(i, j) * c[j];
        c[i] = s / A(i, i);
    }
    std::vector<ld> t1, c1, t2, c2;
    differentiate(t, P, c, t1, c1);
    differentiate(t1, P - 1, c1, t2, c2);
    for (int64_t i = 0; i < n; ++i) {
        if (d1)
            d1[i] = (double)spline_eval(t1, P - 1, c1, x[i]);
        if (d2)
            d2[i] = (double)spline_eval(t2, P - 2, c2, x[i]);
    }
    return true;
}

// numpy.linspace(0, stop, num): i * (stop / (num - 1)), the last node = stop
std::vector<double> linspace0(double stop, int64_t num) {
    std::vector<double> g(num);
    double const step = stop / (double)(num - 1);
    for (int64_t i = 0; i < num; ++i)
        g[i] = (double)i * step;
    g[num - 1] = stop;
    return g;
}

struct NodeSet {
    std::vector<double> f, d1, d2;
};

}  // namespace

struct Setfl {
    struct Element {
        std::string name, lattice;
        double Z = 0, mass = 0, a = 0;
        NodeSet F, rho;
    };
    std::vector<Element> el;
    std::vector<NodeSet> phi;   // by slot j(j+1)/2 + i
    int64_t nrho = 0, nr = 0;
    double drho = 0, dr = 0, rcut_header = 0;
    std::vector<double> r_grid, rho_grid;
};

namespace {

bool split(const std::string &line, std::vector<std::string> &tok) {
    tok.clear();
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && isspace((unsigned char)line[i]))
            ++i;
        size_t const j = i;
        while (i < line.size() && !isspace((unsigned char)line[i]))
            ++i;
        if (i > j)
            tok.push_back(line.substr(j, i - j));
    }
    return !tok.empty();
}

bool to_double(const std::string &s, double &v) {
    errno = 0;
    char *end = nullptr;
    v = strtod(s.c_str(), &end);
    return end && *end == '\0' && end != s.c_str();
}

}  // namespace
}  // namespace mdl

using namespace mdl;

extern "C" {

mdl_status mdl_spline5_node_derivatives(int64_t n, const double *x, const double *y, double *d1,
                                        double *d2) {
    if (n < 6 || !x || !y) {
        set_error("mdl_spline5_node_derivatives: need n >= 6 nodes and x, y (n = %lld)",
                  (long long)n);
        return MDL_ERR_ARG;
    }
    for (int64_t i = 1; i < n; ++i)
        if (!(x[i] > x[i - 1])) {
            set_error("mdl_spline5_node_derivatives: x must increase strictly (at %lld)",
                      (long long)i);
            return MDL_ERR_ARG;
        }
    if (!spline5_derivatives(n, x, y, d1, d2)) {
        set_error("mdl_spline5_node_derivatives: singular collocation system");
        return MDL_ERR_ARG;
    }
    return MDL_OK;
}

mdl_status mdl_setfl_read(const char *path, int32_t with_derivatives, mdl_setfl **out) {
    if (!path || !out) {
        set_error("mdl_setfl_read: null argument");
        return MDL_ERR_ARG;
    }
    *out = nullptr;
    FILE *fh = fopen(path, "r");
    if (!fh) {
        set_error("mdl_setfl_read: cannot open %s", path);
        return MDL_ERR_ARG;
    }
    std::vector<std::string> lines;
    {
        std::string cur;
        char buf[4096];
        while (fgets(buf, sizeof buf, fh)) {
            cur += buf;
            if (!cur.empty() && cur.back() == '\n') {
                cur.pop_back();
                if (!cur.empty() && cur.back() == '\r')
                    cur.pop_back();
                lines.push_back(cur);
                cur.clear();
            }
        }
        if (!cur.empty())
            lines.push_back(cur);
        fclose(fh);
    }
    auto fail = [&](const char *what) {
        set_error("mdl_setfl_read: %s: %s", path, what);
        return MDL_ERR_ARG;
    };
    if (lines.size() < 6)
        return fail("fewer than 6 lines");
    auto S = std::make_unique<Setfl>();
    std::vector<std::string> tok;
    // line 4: Nel and the element names
    if (!split(lines[3], tok))
        return fail("empty element line");
    int const nel = atoi(tok[0].c_str());
    if (nel < 1 || nel > MAX_TYPES || (int)tok.size() < 1 + nel)
        return fail("bad element line (count or names)");
    S->el.resize(nel);
    for (int e = 0; e < nel; ++e)
        S->el[e].name = tok[1 + e];
    // line 5: Nrho drho Nr dr rcut (Angstrom units)
    double v[5];
    if (!split(lines[4], tok) || tok.size() < 5)
        return fail("bad grid line");
    for (int k = 0; k < 5; ++k)
        if (!to_double(tok[k], v[k]))
            return fail("bad number in the grid line");
    S->nrho = (int64_t)v[0];
    S->nr = (int64_t)v[2];
    if (S->nrho < 6 || S->nr < 6 || !(v[1] > 0.) || !(v[3] > 0.))
        return fail("grid needs >= 6 nodes and positive spacings");
    double const A3 = std::pow(ANGSTROM, 3.0);
    S->drho = v[1] / A3;
    S->dr = v[3] * ANGSTROM;
    S->rcut_header = v[4] * ANGSTROM;
    S->r_grid = linspace0(S->dr * ((double)S->nr - 1.0), S->nr);
    S->rho_grid = linspace0(S->drho * ((double)S->nrho - 1.0), S->nrho);

    size_t pos = 5;
    std::string err;
    auto take = [&](int64_t n, std::vector<double> &dst) -> bool {
        dst.clear();
        dst.reserve(n);
        while ((int64_t)dst.size() < n) {
            if (pos >= lines.size()) {
                err = "file ends inside a table";
                return false;
            }
            split(lines[pos], tok);
            ++pos;
            for (auto &s : tok) {
                double x;
                if (!to_double(s, x)) {
                    err = "bad number in a table";
                    return false;
                }
                dst.push_back(x);
            }
        }
        if ((int64_t)dst.size() != n) {
            err = "a table line crosses a block boundary";
            return false;
        }
        return true;
    };
    auto derive = [&](const std::vector<double> &grid, NodeSet &ns) -> bool {
        ns.d1.assign(ns.f.size(), 0.);
        ns.d2.assign(ns.f.size(), 0.);
        if (!with_derivatives)
            return true;
        return spline5_derivatives((int64_t)grid.size(), grid.data(), ns.f.data(), ns.d1.data(),
                                   ns.d2.data());
    };
    for (int e = 0; e < nel; ++e) {
        if (pos >= lines.size() || !split(lines[pos], tok) || tok.size() < 3)
            return fail("missing element header line");
        ++pos;
        auto &E = S->el[e];
        if (!to_double(tok[0], E.Z) || !to_double(tok[1], E.mass) || !to_double(tok[2], E.a))
            return fail("bad element header line");
        if (tok.size() > 3)
            E.lattice = tok[3];
        if (!take(S->nrho, E.F.f) || !take(S->nr, E.rho.f))
            return fail(err.c_str());
        for (auto &x : E.F.f)
            x *= EV;
        for (auto &x : E.rho.f)
            x /= A3;
        if (!derive(S->r_grid, E.rho) || !derive(S->rho_grid, E.F))
            return fail("spline derivatives failed");
    }
    int const npair = nel * (nel + 1) / 2;
    S->phi.resize(npair);
    int i = 0, j = 0;
    for (int b = 0; b < npair; ++b) {
        std::vector<double> phir;
        if (!take(S->nr, phir))
            return fail(err.c_str());
        NodeSet &ns = S->phi[j * (j + 1) / 2 + i];
        ns.f.resize(S->nr);
        // phi = (r phi) / r * eV * A, phi[0] = phi[1] (get_potential.py:127-128)
        for (int64_t k = 1; k < S->nr; ++k)
            ns.f[k] = phir[k] / S->r_grid[k] * EV * ANGSTROM;
        ns.f[0] = ns.f[1];
        if (!derive(S->r_grid, ns))
            return fail("spline derivatives failed");
        // blocks 11, 21, 22, 31, ...: (i, j) = (0,0), (0,1), (1,1), (0,2), ...
        if (i == j) {
            ++j;
            i = 0;
        } else {
            ++i;
        }
    }
    *out = reinterpret_cast<mdl_setfl *>(S.release());
    return MDL_OK;
}

mdl_status mdl_setfl_destroy(mdl_setfl *s) {
    delete reinterpret_cast<Setfl *>(s);
    return MDL_OK;
}

mdl_status mdl_setfl_grid(const mdl_setfl *s_, int32_t *n_elements, int64_t *n_rho, double *drho,
                          int64_t *n_r, double *dr, double *rcut, double *rhomax,
                          double *rcut_header) {
    const Setfl *s = reinterpret_cast<const Setfl *>(s_);
    if (!s) {
        set_error("mdl_setfl_grid: null handle");
        return MDL_ERR_ARG;
    }
    if (n_elements)
        *n_elements = (int32_t)s->el.size();
    if (n_rho)
        *n_rho = s->nrho;
    if (drho)
        *drho = s->drho;
    if (n_r)
        *n_r = s->nr;
    if (dr)
        *dr = s->dr;
    if (rcut)
        *rcut = s->r_grid.back();
    if (rhomax)
        *rhomax = s->rho_grid.back();
    if (rcut_header)
        *rcut_header = s->rcut_header;
    return MDL_OK;
}

mdl_status mdl_setfl_element(const mdl_setfl *s_, int32_t e, char *name, int32_t name_len,
                             double *Z, double *mass, double *a_lattice, char *lattice,
                             int32_t lattice_len) {
    const Setfl *s = reinterpret_cast<const Setfl *>(s_);
    if (!s || e < 0 || e >= (int32_t)s->el.size()) {
        set_error("mdl_setfl_element: bad handle or element index %d", e);
        return MDL_ERR_ARG;
    }
    auto const &E = s->el[e];
    if (name && name_len > 0)
        snprintf(name, (size_t)name_len, "%s", E.name.c_str());
    if (lattice && lattice_len > 0)
        snprintf(lattice, (size_t)lattice_len, "%s", E.lattice.c_str());
    if (Z)
        *Z = E.Z;
    if (mass)
        *mass = E.mass;
    if (a_lattice)
        *a_lattice = E.a;
    return MDL_OK;
}

mdl_status mdl_setfl_nodes(const mdl_setfl *s_, int32_t which, int32_t index, double *f,
                           double *d1, double *d2) {
    const Setfl *s = reinterpret_cast<const Setfl *>(s_);
    if (!s) {
        set_error("mdl_setfl_nodes: null handle");
        return MDL_ERR_ARG;
    }
    const NodeSet *ns = nullptr;
    if ((which == MDL_SETFL_F || which == MDL_SETFL_RHO) && index >= 0 &&
        index < (int32_t)s->el.size())
        ns = which == MDL_SETFL_F ? &s->el[index].F : &s->el[index].rho;
    else if (which == MDL_SETFL_PHI && index >= 0 && index < (int32_t)s->phi.size())
        ns = &s->phi[index];
    if (!ns) {
        set_error("mdl_setfl_nodes: bad array %d or index %d", which, index);
        return MDL_ERR_ARG;
    }
    size_t const n = ns->f.size();
    if (f)
        memcpy(f, ns->f.data(), n * sizeof(double));
    if (d1)
        memcpy(d1, ns->d1.data(), n * sizeof(double));
    if (d2)
        memcpy(d2, ns->d2.data(), n * sizeof(double));
    return MDL_OK;
}

}  // extern "C"
