// Potential tables: host construction of the per-interval polynomial records
// and the device evaluation entry point (mdl_table_*).
//
// Quintic tables follow poly_table.hpp:33-115 of the reference: on every
// interval the quintic Hermite polynomial fixed by f, f', f'' at both nodes,
// coefficients in t = (x - x_i)/h formed in long double (poly_table.hpp:45-62).
//
// B-spline tables follow poly_table.hpp:120-135: three independent cardinal
// cubic B-splines (Boost.Math cardinal_cubic_b_spline, which the reference
// calls; the library is not part of this build) through the node values of f,
// f' and f''. The spline is the interpolant s(x) = avg + sum_k beta_k B3(x/h+1-k)
// of y_k = f_k - avg with end slopes from the one-sided five-point O(h^4)
// formulas; here each interval's part of it is stored as a cubic in t so the
// device evaluates three Horner cubics instead of 3 x 4 basis functions.
#include "mdl_internal.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>

namespace mdl {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

mdl_status hip_fail(hipError_t e, const char *what) {
    set_error("HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
    return MDL_ERR_DEVICE;
}

// Quintic Hermite coefficients of one interval (poly_table.hpp:45-62).
static void quintic_record(double h, double fa, double fb, double f1a, double f1b, double f2a,
                           double f2b, double *c) {
    using ld = long double;
    ld const hl = (ld)h;
    ld const dy = (ld)fb - (ld)fa;
    ld const D0 = hl * (ld)f1a, D1 = hl * (ld)f1b;
    ld const S0 = hl * hl * (ld)f2a, S1 = hl * hl * (ld)f2b;
    c[0] = fa;
    c[1] = (double)D0;
    c[2] = (double)(0.5L * S0);
    c[3] = (double)(10.L * dy - 6.L * D0 - 4.L * D1 - 0.5L * (3.L * S0 - S1));
    c[4] = (double)(-15.L * dy + 8.L * D0 + 7.L * D1 + 0.5L * (3.L * S0 - 2.L * S1));
    c[5] = (double)(6.L * dy - 3.L * (D0 + D1) - 0.5L * (S0 - S1));
}

// Coefficients beta (size n+2) of the cardinal cubic B-spline through n >= 5
// values, plus the subtracted mean; long double Thomas solve.
static bool bspline_beta(double h, int64_t n, const double *f, std::vector<long double> &beta,
                         long double &avg) {
    using ld = long double;
    if (n < 5)
        return false;
    avg = 0;
    for (int64_t k = 0; k < n; ++k)
        avg += f[k];
    avg /= (ld)n;
    ld const hl = h;
    ld const a1 = (-25.L * f[0] + 48.L * f[1] - 36.L * f[2] + 16.L * f[3] - 3.L * f[4]) / (12.L * hl);
    ld const b1 = (25.L * f[n - 1] - 48.L * f[n - 2] + 36.L * f[n - 3] - 16.L * f[n - 4] +
                   3.L * f[n - 5]) / (12.L * hl);
    std::vector<ld> cp(n), dp(n);
    // row 0: 4 b0 + 2 b1 = 6 y0 + 2 h a1; rows k: b_{k-1} + 4 b_k + b_{k+1} = 6 y_k;
    // row n-1: 2 b_{n-2} + 4 b_{n-1} = 6 y_{n-1} - 2 h b1
    cp[0] = 2.L / 4.L;
    dp[0] = (6.L * ((ld)f[0] - avg) + 2.L * hl * a1) / 4.L;
    for (int64_t k = 1; k < n; ++k) {
        ld const a = (k == n - 1) ? 2.L : 1.L;
        ld const c = (k == n - 1) ? 0.L : 1.L;
        ld const d = 6.L * ((ld)f[k] - avg) - ((k == n - 1) ? 2.L * hl * b1 : 0.L);
        ld const m = 4.L - a * cp[k - 1];
        cp[k] = c / m;
        dp[k] = (d - a * dp[k - 1]) / m;
    }
    beta.assign(n + 2, 0.L);
    ld *x = beta.data() + 1;
    x[n - 1] = dp[n - 1];
    for (int64_t k = n - 2; k >= 0; --k)
        x[k] = dp[k] - cp[k] * x[k + 1];
    beta[0] = x[1] - 2.L * hl * a1;
    beta[n + 1] = x[n - 2] + 2.L * hl * b1;
    return true;
}

// The spline restricted to interval i as a cubic in t: the four basis
// functions covering [x_i, x_{i+1}] are (1-t)^3/6, (4-6t^2+3t^3)/6,
// (1+3t+3t^2-3t^3)/6 and t^3/6, with coefficients beta_i .. beta_{i+3}.
static void bspline_cubic(const std::vector<long double> &beta, long double avg, int64_t i,
                          double *c) {
    long double const b0 = beta[i], b1 = beta[i + 1], b2 = beta[i + 2], b3 = beta[i + 3];
    c[0] = (double)(avg + (b0 + 4.L * b1 + b2) / 6.L);
    c[1] = (double)((b2 - b0) / 2.L);
    c[2] = (double)((b0 - 2.L * b1 + b2) / 2.L);
    c[3] = (double)((-b0 + 3.L * b1 - 3.L * b2 + b3) / 6.L);
}

__global__ void k_table_eval(int kind, const double *rec, int n, double inv_h, double inv_h2,
                             int64_t m, const double *x, double *v, double *d1, double *d2) {
    int64_t const k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m)
        return;
    if (kind == MDL_MORSE || kind == MDL_LJ) {
        double V, dV, d2V;
        if (kind == MDL_MORSE)
            pair_analytic<MDL_MORSE>(rec, x[k], V, dV, d2V);
        else
            pair_analytic<MDL_LJ>(rec, x[k], V, dV, d2V);
        v[k] = V;
        d1[k] = dV;
        d2[k] = d2V;
        return;
    }
    double t;
    int const i = tab_locate(x[k], inv_h, n, t);
    if (kind == MDL_QUINTIC) {
        Poly<MDL_QUINTIC> p;
        p.load(rec + (size_t)i * 6);
        v[k] = p.v(t);
        d1[k] = p.d1(t, inv_h);
        d2[k] = p.d2(t, inv_h2);
    } else {
        Poly<MDL_BSPLINE> p;
        p.load(rec + (size_t)i * 12);
        v[k] = p.v(t);
        d1[k] = p.d1(t, inv_h);
        d2[k] = p.d2(t, inv_h2);
    }
}

}  // namespace mdl

using namespace mdl;

extern "C" {

const char *mdl_last_error(void) { return g_err; }

mdl_status mdl_device_info(char *name, int32_t name_len, int32_t *compute_units) {
    int dev = 0;
    MDL_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MDL_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0)
        snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    if (compute_units)
        *compute_units = prop.multiProcessorCount;
    return MDL_OK;
}

mdl_status mdl_table_create(int32_t kind, double h, int64_t n_nodes, const double *f,
                            const double *f1, const double *f2, mdl_table **out) {
    if (!out || !f || !f1 || !f2) {
        set_error("mdl_table_create: null argument");
        return MDL_ERR_ARG;
    }
    *out = nullptr;
    // poly_table.hpp:39-44
    if (n_nodes < 2 || !(h > 0.)) {
        set_error("FATAL: QuinticTable needs >= 2 nodes, equal array sizes, and h > 0 "
                  "(nodes %lld, h %g)", (long long)n_nodes, h);
        return MDL_ERR_TABLE;
    }
    if (kind != MDL_QUINTIC && kind != MDL_BSPLINE) {
        set_error("FATAL: unknown table interpolation %d (use quintic or bspline)", kind);
        return MDL_ERR_TABLE;
    }
    auto t = std::make_unique<HostTable>();
    t->kind = kind;
    t->h = h;
    t->inv_h = 1. / h;
    t->inv_h2 = 1. / (h * h);
    t->n = n_nodes - 1;
    if (kind == MDL_QUINTIC) {
        t->stride = 6;
        t->rec.resize((size_t)t->n * 6);
        for (int64_t i = 0; i < t->n; ++i)
            quintic_record(h, f[i], f[i + 1], f1[i], f1[i + 1], f2[i], f2[i + 1],
                           t->rec.data() + 6 * i);
    } else {
        t->stride = 12;
        t->rec.resize((size_t)t->n * 12);
        const double *src[3] = {f, f1, f2};
        for (int s = 0; s < 3; ++s) {
            std::vector<long double> beta;
            long double avg;
            if (!bspline_beta(h, n_nodes, src[s], beta, avg)) {
                set_error("a cubic B-spline table needs at least 5 nodes (got %lld)",
                          (long long)n_nodes);
                return MDL_ERR_TABLE;
            }
            for (int64_t i = 0; i < t->n; ++i)
                bspline_cubic(beta, avg, i, t->rec.data() + 12 * i + 4 * s);
        }
    }
    *out = reinterpret_cast<mdl_table *>(t.release());
    return MDL_OK;
}

mdl_status mdl_table_create_analytic(int32_t kind, const double *param, mdl_table **out) {
    if (!out || !param) {
        set_error("mdl_table_create_analytic: null argument");
        return MDL_ERR_ARG;
    }
    *out = nullptr;
    if (kind != MDL_MORSE && kind != MDL_LJ) {
        set_error("mdl_table_create_analytic: unknown analytic form %d (use Morse or LJ)", kind);
        return MDL_ERR_ARG;
    }
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(param[k])) {
            set_error("mdl_table_create_analytic: parameter %d is not finite", k);
            return MDL_ERR_ARG;
        }
    if (kind == MDL_LJ && !(param[2] > 0.)) {
        set_error("mdl_table_create_analytic: Lennard-Jones rcut must be > 0 (got %g)", param[2]);
        return MDL_ERR_ARG;
    }
    auto t = std::make_unique<HostTable>();
    t->kind = kind;
    t->n = 1;   // one record: the parameters
    t->stride = PAIR_PARAMS;
    t->rec.assign(PAIR_PARAMS, 0.);
    t->rec[0] = param[0];
    t->rec[1] = param[1];
    if (kind == MDL_MORSE) {
        t->rec[2] = param[2];
    } else {
        // vcut, potential.hpp:57-59: 4 epsilon ((s/rc)^12 - (s/rc)^6)
        double const v6 = pow6_binary(param[1] / param[2]);
        t->rec[2] = 4. * param[0] * (v6 * v6 - v6);
    }
    *out = reinterpret_cast<mdl_table *>(t.release());
    return MDL_OK;
}

mdl_status mdl_table_destroy(mdl_table *tab) {
    HostTable *t = reinterpret_cast<HostTable *>(tab);
    if (!t)
        return MDL_OK;
    if (t->d_rec)
        (void)hipFree(t->d_rec);
    delete t;
    return MDL_OK;
}

mdl_status mdl_table_coefficients(const mdl_table *tab, double *out, int64_t *n_intervals,
                                  int32_t *stride) {
    const HostTable *t = reinterpret_cast<const HostTable *>(tab);
    if (!t) {
        set_error("mdl_table_coefficients: null table");
        return MDL_ERR_ARG;
    }
    if (n_intervals)
        *n_intervals = t->n;
    if (stride)
        *stride = t->stride;
    if (out)
        memcpy(out, t->rec.data(), sizeof(double) * t->rec.size());
    return MDL_OK;
}

mdl_status mdl_table_eval(mdl_table *tab, int64_t n, const double *x, double *value, double *d1,
                          double *d2) {
    HostTable *t = reinterpret_cast<HostTable *>(tab);
    if (!t || n < 0 || (n > 0 && !x)) {
        set_error("mdl_table_eval: bad argument");
        return MDL_ERR_ARG;
    }
    if (n == 0)
        return MDL_OK;
    int dev = 0;
    MDL_HIP(hipGetDevice(&dev));
    if (!t->d_rec || t->device != dev) {
        MDL_HIP(hipMalloc(&t->d_rec, sizeof(double) * t->rec.size()));
        MDL_HIP(hipMemcpy(t->d_rec, t->rec.data(), sizeof(double) * t->rec.size(),
                          hipMemcpyHostToDevice));
        t->device = dev;
    }
    double *buf = nullptr;
    MDL_HIP(hipMalloc(&buf, sizeof(double) * 4 * (size_t)n));
    double *dx = buf, *dv = buf + n, *d1v = buf + 2 * n, *d2v = buf + 3 * n;
    hipError_t e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        unsigned const blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_table_eval, dim3(blocks), dim3(256), 0, 0, t->kind, t->d_rec,
                           (int)t->n, t->inv_h, t->inv_h2, n, dx, dv, d1v, d2v);
        e = hipGetLastError();
    }
    if (e == hipSuccess && value)
        e = hipMemcpy(value, dv, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && d1)
        e = hipMemcpy(d1, d1v, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && d2)
        e = hipMemcpy(d2, d2v, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return hip_fail(e, "mdl_table_eval");
    return MDL_OK;
}

}  // extern "C"
