// MDDriver on the GPU: creation (md_driver.cpp:16-160), integrate()
// (md_driver.cpp:563-601), calculate_neighbors() (md_driver.cpp:279-298) and
// the attribute getters of md_driver_module.cpp:103-140, as the mdl_driver_*
// C-ABI of include/mdelast.h.
#include "mdl_kernels.hpp"
#include "mdl_mixed.hpp"
#include "mdl_cv.hpp"
#include "../../include/mdelast_cv.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <tuple>

namespace mdl {

constexpr int GRID_MAX = PART_STRIDE;   // blocks of the pair passes (multiple of 8)
constexpr int GRID_ATOM_MAX = 1024;     // blocks of the kick kernels

using DensFn = void (*)(AtomArrays, ListArrays, RTables, FTables, int, int, int, double, double,
                        DevState *, double *);
using ForceFn = void (*)(AtomArrays, ListArrays, RTables, int, int, double, double, double,
                         double, DevState *, double *);

template <int KIND, bool S> static DensFn dens_fn(int lpa) {
    switch (lpa) {
    case 4: return k_density<KIND, 4, S>;
    case 8: return k_density<KIND, 8, S>;
    case 16: return k_density<KIND, 16, S>;
    case 32: return k_density<KIND, 32, S>;
    default: return k_density<KIND, 64, S>;
    }
}
template <int KIND, bool S, bool E> static ForceFn force_fn(int lpa) {
    switch (lpa) {
    case 4: return k_force<KIND, 4, S, E>;
    case 8: return k_force<KIND, 8, S, E>;
    case 16: return k_force<KIND, 16, S, E>;
    case 32: return k_force<KIND, 32, S, E>;
    default: return k_force<KIND, 64, S, E>;
    }
}
using DensMxFn = void (*)(AtomArrays, ListArrays, RTablesF, FTablesF, int, int, int, double,
                          DevState *, double *);
using ForceMxFn = void (*)(AtomArrays, ListArrays, RTablesF, int, int, float, float, DevState *,
                           double *);
template <int KIND, bool S> static DensMxFn dens_mx_fn(int lpa) {
    switch (lpa) {
    case 4: return k_density_mx<KIND, 4, S>;
    case 8: return k_density_mx<KIND, 8, S>;
    case 16: return k_density_mx<KIND, 16, S>;
    case 32: return k_density_mx<KIND, 32, S>;
    default: return k_density_mx<KIND, 64, S>;
    }
}
template <int KIND, bool S> static ForceMxFn force_mx_fn(int lpa) {
    switch (lpa) {
    case 4: return k_force_mx<KIND, 4, S>;
    case 8: return k_force_mx<KIND, 8, S>;
    case 16: return k_force_mx<KIND, 16, S>;
    case 32: return k_force_mx<KIND, 32, S>;
    default: return k_force_mx<KIND, 64, S>;
    }
}
static DensMxFn pick_dens_mx(int kind, int lpa, bool s) {
    if (kind == MDL_QUINTIC)
        return s ? dens_mx_fn<MDL_QUINTIC, true>(lpa) : dens_mx_fn<MDL_QUINTIC, false>(lpa);
    return s ? dens_mx_fn<MDL_BSPLINE, true>(lpa) : dens_mx_fn<MDL_BSPLINE, false>(lpa);
}
static ForceMxFn pick_force_mx(int kind, int lpa, bool s) {
    if (kind == MDL_QUINTIC)
        return s ? force_mx_fn<MDL_QUINTIC, true>(lpa) : force_mx_fn<MDL_QUINTIC, false>(lpa);
    return s ? force_mx_fn<MDL_BSPLINE, true>(lpa) : force_mx_fn<MDL_BSPLINE, false>(lpa);
}
static DensFn pick_dens(int kind, int lpa, bool s) {
    if (kind == MDL_QUINTIC)
        return s ? dens_fn<MDL_QUINTIC, true>(lpa) : dens_fn<MDL_QUINTIC, false>(lpa);
    return s ? dens_fn<MDL_BSPLINE, true>(lpa) : dens_fn<MDL_BSPLINE, false>(lpa);
}
static ForceFn pick_force(int kind, int lpa, bool s, bool e) {
    // analytic pair forms run without EAM (checked at creation)
    if (kind == MDL_MORSE)
        return s ? force_fn<MDL_MORSE, true, false>(lpa) : force_fn<MDL_MORSE, false, false>(lpa);
    if (kind == MDL_LJ)
        return s ? force_fn<MDL_LJ, true, false>(lpa) : force_fn<MDL_LJ, false, false>(lpa);
    if (kind == MDL_QUINTIC) {
        if (e)
            return s ? force_fn<MDL_QUINTIC, true, true>(lpa) : force_fn<MDL_QUINTIC, false, true>(lpa);
        return s ? force_fn<MDL_QUINTIC, true, false>(lpa) : force_fn<MDL_QUINTIC, false, false>(lpa);
    }
    if (e)
        return s ? force_fn<MDL_BSPLINE, true, true>(lpa) : force_fn<MDL_BSPLINE, false, true>(lpa);
    return s ? force_fn<MDL_BSPLINE, true, false>(lpa) : force_fn<MDL_BSPLINE, false, false>(lpa);
}

// Autotuner cache: (device, N, ntypes, kind, precision, EAM or not, list stride
// bucket) -> lanes per atom; drivers may be created from several threads
using LpaKey = std::tuple<int, int64_t, int, int, int, int, int>;
static std::map<LpaKey, int> g_lpa_cache;
static std::mutex g_lpa_mutex;

// The batched replica runs a driver leads (comes first in the list of): the
// launch tables in one device buffer, uploaded once and sent again only when
// they change, and the captured graph of one chunk of batched steps.
struct RepSet {
    void *buf = nullptr;
    size_t cap = 0;
    std::vector<char> image;       // what buf holds
    std::vector<int64_t> shape;    // launch sizes and pass groups of that image
    hipGraphExec_t exec = nullptr;
    int64_t stride = 0, rebuild = 0, steps = 0, nrec = 0;

    void drop_graph() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    void release() {
        drop_graph();
        if (buf)
            (void)hipFree(buf);
        buf = nullptr;
        cap = 0;
        image.clear();
        shape.clear();
    }
};

struct Driver {
    int device = 0;
    hipStream_t stream = nullptr;
    int N = 0, ntypes = 0, npair = 0, kind = MDL_QUINTIC, has_eam = 1, precision = MDL_FP64;
    int lpa = 16;
    bool use_cells = false;
    int sample = 1;
    int64_t cycle_host = 0;
    bool build_pending = true;     // calculate_neighbors() or the constructor asked for a build
    bool failed = false;
    AtomArrays a{};
    ListArrays l{};
    RTables rt{};
    FTables ft{};
    RTablesF rtf{};
    FTablesF ftf{};
    float *d_rrec_f = nullptr, *d_frec_f = nullptr, *d_prec_f = nullptr;
    double rc2 = 0, rc_pair = 0, rc_rho = 0, rmin = 0;
    double *d_rrec = nullptr, *d_frec = nullptr;
    double *pA = nullptr, *pB = nullptr, *pK = nullptr;
    int gridA = 8, gridB = 8, gridK = 1;
    double *d_stage = nullptr;
    double *d_totals = nullptr, *d_dmax = nullptr;
    DevState *gs = nullptr;
    DevState hs{};                 // host mirror (refreshed by sync_state)
    std::vector<double> init_pos;  // positions attribute before the first step
    std::vector<void *> allocs;
    // timing of the pair passes
    std::vector<hipEvent_t> ev;
    hipEvent_t ev_cv[2] = {nullptr, nullptr};   // ordering with a caller's stream
    std::vector<int64_t> type_count;             // atoms per species
    // control-variate inputs of the sampled steps of a run, written into frame
    // slots in turn (mdl_driver_set_cv_capture)
    struct CvCapture {
        float *D = nullptr, *J = nullptr, *L = nullptr;
        double *pos = nullptr, *force = nullptr;
        int *nbr = nullptr, *ncnt = nullptr;   // optional pair lists (mdl_driver_set_cv_capture_pairs)
        int width = 0;
        int slots = 0, next = 0, list = 0;   // list: pair_list's choice
        CvBasis basis{};
    } cvc;
    // run() replays graph_steps consecutive steps as one captured graph (0: off)
    int graph_steps = 0;
    hipGraphExec_t gexec = nullptr;
    int64_t g_stride = 0, g_rebuild = 0, g_nrec = 0, g_cap = 0;
    double *g_rec = nullptr;
    uint64_t uid = 0;   // identifies the driver in the joint-graph cache
    // device staging of run()'s records, kept between calls (grown on demand)
    double *run_rec = nullptr;
    int64_t run_rec_cap = 0;
    RepSet rep;
    double t_dens = 0, t_force = 0, t_step = 0;

    template <class T> mdl_status alloc(T **p, size_t n) {
        void *q = nullptr;
        MDL_HIP(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
        allocs.push_back(q);
        *p = static_cast<T *>(q);
        return MDL_OK;
    }
    ~Driver() {
        for (auto e : ev)
            (void)hipEventDestroy(e);
        for (auto e : ev_cv)
            if (e)
                (void)hipEventDestroy(e);
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        rep.release();
        for (void *p : allocs)
            (void)hipFree(p);
        if (stream)
            (void)hipStreamDestroy(stream);
    }

    int grid_pairs(int lanes) const {
        int const apb = BLOCK / lanes;
        int g = (N + apb - 1) / apb;
        g = std::min(std::max(g, 1), GRID_MAX);
        return (g + 7) / 8 * 8;
    }

    int gridL = 1;                  // blocks of k_list: no more than the device holds at once
    int gridL_solo = 1;             // ... when this driver runs alone
    int64_t list_cap = 1;           // k_list blocks the device holds at once
    unsigned *d_bar = nullptr, *d_ticket_fin = nullptr;

    // The per-step list check (and the rare build it finds due).
    void launch_list() {
        k_list<<<gridL, BLOCK, 0, stream>>>(a, l, N, gs, d_bar, d_dmax, gridK, use_cells ? 1 : 0);
    }

    // A requested build, as a chain of kernels.
    void launch_build() {
        int const gN = (N + BLOCK - 1) / BLOCK;
        if (use_cells) {
            k_rb_bin<<<gN, BLOCK, 0, stream>>>(a, l, N, gs);
            k_rb_scan<<<1, BLOCK, 0, stream>>>(l, gs);
            k_rb_scatter<<<gN, BLOCK, 0, stream>>>(l, N, gs);
            k_rb_sort<<<l.ncells, BLOCK, 0, stream>>>(l, gs);
            k_rb_build<<<l.ncells, BLOCK, 0, stream>>>(a, l, N, gs, 1);
        } else {
            k_rb_begin<<<1, 64, 0, stream>>>(gs);
            int const gB = (int)(((int64_t)N * LIST_LANES + BLOCK - 1) / BLOCK);
            k_rb_build<<<gB, BLOCK, 0, stream>>>(a, l, N, gs, 0);
        }
    }

    void launch_finish(int mode, double *rec_out, int next_sample) {
        k_finish<<<NTOTAL, BLOCK, 0, stream>>>(gs, pA, gridA, pB, gridB, pK, gridK, d_totals,
                                               d_ticket_fin, mode, rec_out, next_sample);
    }

    void launch_forces(bool s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                       hipEvent_t e2 = nullptr) {
        if (e0)
            (void)hipEventRecord(e0, stream);
        if (precision == MDL_MIXED) {
            double const rc2A = rc2 * 1e16;
            DensMxFn fa = pick_dens_mx(kind, lpa, s);
            fa<<<gridA, BLOCK, 0, stream>>>(a, l, rtf, ftf, N, npair, ntypes, rc2A, gs, pA);
            if (e1)
                (void)hipEventRecord(e1, stream);
            ForceMxFn fb = pick_force_mx(kind, lpa, s);
            fb<<<gridB, BLOCK, 0, stream>>>(a, l, rtf, N, npair, (float)(rmin * 1e8),
                                            (float)rc2A, gs, pB);
            if (e2)
                (void)hipEventRecord(e2, stream);
            return;
        }
        if (has_eam) {
            DensFn fa = pick_dens(kind, lpa, s);
            fa<<<gridA, BLOCK, 0, stream>>>(a, l, rt, ft, N, npair, ntypes, rc2, rc_rho, gs, pA);
        }
        if (e1)
            (void)hipEventRecord(e1, stream);
        ForceFn fb = pick_force(kind, lpa, s, has_eam != 0);
        fb<<<gridB, BLOCK, 0, stream>>>(a, l, rt, N, npair, rc2, rc_pair, rc_rho, rmin, gs, pB);
        if (e2)
            (void)hipEventRecord(e2, stream);
    }

    // One step after its first half (k_pre or the previous k_finish) has run.
    // build: rebuild the list after the drift (else only check it).
    // next_sample < 0: end with the post step only; otherwise also run the next
    // step's first half with that sampling flag.
    hipError_t launch_step(bool s, double *rec_out, hipEvent_t *tev, bool build,
                           int next_sample) {
        k_kick_drift<<<gridK, BLOCK, 0, stream>>>(a, N, gs, d_dmax);
        if (build)
            launch_build();
        else
            launch_list();
        if (tev)
            launch_forces(s, tev[0], tev[1], tev[2]);
        else
            launch_forces(s);
        if (s)
            k_kick2<true><<<gridK, BLOCK, 0, stream>>>(a, N, gs, 0, has_eam, pK);
        else
            k_kick2<false><<<gridK, BLOCK, 0, stream>>>(a, N, gs, 0, has_eam, pK);
        launch_finish(1, rec_out, next_sample);
        ++cycle_host;
        return hipGetLastError();
    }

    void drop_graph() {
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
    }

    // Records of one captured chunk: count them and make room in g_rec.
    mdl_status chunk_records(int64_t stride, int64_t &nrec) {
        nrec = 0;
        for (int64_t j = 0; j < graph_steps; ++j)
            if ((cycle_host + j + 1) % stride == 0)
                ++nrec;
        if (nrec > g_cap) {
            drop_graph();   // it writes to the previous buffer
            if (mdl_status st = alloc(&g_rec, (size_t)(MDL_NRECORD * nrec)))
                return st;
            g_cap = nrec;
        }
        return MDL_OK;
    }

    // The launches of graph_steps steps from the current cycle (a multiple of
    // graph_steps) with the given sampling stride and build interval, as
    // launch_step would issue them, every step continuing into the next one;
    // sampled records go to g_rec in order. Used under stream capture: the
    // cycle counter is restored.
    hipError_t chunk_launches(int64_t stride, int64_t rebuild) {
        int64_t const c0 = cycle_host;
        hipError_t e = hipSuccess;
        int64_t irec = 0;
        for (int64_t j = 0; j < graph_steps && e == hipSuccess; ++j) {
            int64_t const cyc = c0 + j;
            bool const smp = (cyc + 1) % stride == 0;
            bool const build = rebuild > 0 && cyc % rebuild == 0;
            double *ro = smp ? g_rec + MDL_NRECORD * irec++ : nullptr;
            e = launch_step(smp, ro, nullptr, build, (cyc + 2) % stride == 0 ? 1 : 0);
        }
        cycle_host = c0;
        return e;
    }

    // Capture one chunk of this driver alone as gexec.
    mdl_status capture_graph(int64_t stride, int64_t rebuild) {
        drop_graph();
        int64_t nrec = 0;
        if (mdl_status st = chunk_records(stride, nrec))
            return st;
        hipGraph_t graph = nullptr;
        MDL_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = chunk_launches(stride, rebuild);
        hipError_t const e2 = hipStreamEndCapture(stream, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph)
                (void)hipGraphDestroy(graph);
            return hip_fail(e != hipSuccess ? e : e2, "step graph capture");
        }
        e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            gexec = nullptr;
            return hip_fail(e, "step graph instantiation");
        }
        g_stride = stride;
        g_rebuild = rebuild;
        g_nrec = nrec;
        return MDL_OK;
    }

    mdl_status sync_state() {
        MDL_HIP(hipMemcpyAsync(&hs, gs, sizeof(DevState), hipMemcpyDeviceToHost, stream));
        MDL_HIP(hipStreamSynchronize(stream));
        return MDL_OK;
    }

    // Translate device error bits into the reference's FATAL messages.
    mdl_status check_errors() {
        int const e = hs.err;
        if (!e)
            return MDL_OK;
        failed = true;
        if (e & DE_RHOMAX) {
            set_error("FATAL: EAM F rho >rhomax_fatal %g %g", hs.err_val[0], hs.err_val[1]);
            return MDL_ERR_RHOMAX;
        }
        if (e & DE_VOLUME) {
            set_error("FATAL: volume change too large... expe=%g Vold=%g Vnew=%g", hs.err_val[0],
                      hs.err_val[1], hs.err_val[2]);
            return MDL_ERR_VOLUME;
        }
        if (e & DE_NBR_OVERFLOW) {
            set_error("neighbour list overflow: an atom has %lld neighbours within r_list, the "
                      "list holds %d", hs.nbr_needed, l.stride);
            return MDL_ERR_NEIGHBORS;
        }
        if (e & DE_CELL_OVERFLOW) {
            set_error("neighbour cells: %.0f atoms fall into one cell (at most %d): the "
                      "positions are non-finite or have collapsed, so the integration has "
                      "diverged", hs.err_val[1], CELL_MAX);
            return MDL_ERR_NEIGHBORS;
        }
        if (e & DE_CELL_GRID) {
            set_error("the box shrank below the cell grid of the neighbour list (L = %g)",
                      hs.err_val[0]);
            return MDL_ERR_NEIGHBORS;
        }
        if (e & DE_MOMENTUM) {
            set_error("FATAL: total momentum = (%g, %g, %g) scale=%g", hs.err_val[0],
                      hs.err_val[1], hs.err_val[2], hs.err_val[3]);
            return MDL_ERR_MOMENTUM;
        }
        if (e & DE_PRESSURE) {
            set_error("FATAL PRESS err=%g %1.16E %1.16E", hs.err_val[0], hs.err_val[1],
                      hs.err_val[2]);
            return MDL_ERR_PRESSURE;
        }
        if (e & DE_CONSERVATION) {
            set_error("FATAL: bad energy conservation... (err %g, current %g)", hs.err_val[0],
                      hs.err_val[1]);
            return MDL_ERR_CONSERVATION;
        }
        set_error("device error bits 0x%x", e);
        return MDL_ERR_DEVICE;
    }
};

static mdl_status bad(const char *msg) {
    set_error("%s", msg);
    return MDL_ERR_ARG;
}

// Displace every atom by up to `amp` (scaled units, per axis) with a fixed
// hash of its index: a thermally disordered stand-in for the autotune, whose
// lattice-perfect start would favour the table-gather pattern of equal
// distances. The caller restores the positions afterwards.
__global__ __launch_bounds__(BLOCK) void k_autotune_jitter(double4 *__restrict__ pos, int N,
                                                           double ax, double ay, double az) {
    int const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= N)
        return;
    unsigned h = (unsigned)i * 2654435761u + 0x9e3779b9u;
    double u[3];
    for (int c = 0; c < 3; ++c) {
        h ^= h >> 16;
        h *= 0x7feb352du;
        h ^= h >> 15;
        h *= 0x846ca68bu;
        h ^= h >> 16;
        u[c] = (double)h * (2.0 / 4294967296.0) - 1.0;   // [-1, 1)
    }
    double4 p = pos[i];
    p.x += ax * u[0];
    p.y += ay * u[1];
    p.z += az * u[2];
    pos[i] = p;
}

// Time the two passes for each candidate lanes-per-atom and keep the fastest,
// on the positions displaced by up to 0.1 A per axis (about the thermal
// amplitude of the alloys here at a few hundred K; the list radius exceeds r_c
// by far more, and the list stays the one built on the undisplaced positions).
// The positions are restored exactly and the error flags cleared afterwards.
static mdl_status autotune(Driver &d) {
    LpaKey const key(d.device, (int64_t)d.N, d.ntypes, d.kind, d.precision, d.has_eam,
                     d.l.stride / 32);
    {
        std::lock_guard<std::mutex> lock(g_lpa_mutex);
        auto it = g_lpa_cache.find(key);
        if (it != g_lpa_cache.end()) {
            d.lpa = it->second;
            d.gridA = d.gridB = d.grid_pairs(d.lpa);
            return MDL_OK;
        }
    }
    hipEvent_t e0, e1;
    MDL_HIP(hipEventCreate(&e0));
    MDL_HIP(hipEventCreate(&e1));
    double4 *saved_pos = nullptr;
    MDL_HIP(hipMalloc(&saved_pos, sizeof(double4) * (size_t)d.N));
    MDL_HIP(hipMemcpyAsync(saved_pos, d.a.pos, sizeof(double4) * (size_t)d.N,
                           hipMemcpyDeviceToDevice, d.stream));
    {
        double const amp = 0.1e-8;   // 0.1 A in cm
        k_autotune_jitter<<<(unsigned)((d.N + BLOCK - 1) / BLOCK), BLOCK, 0, d.stream>>>(
            d.a.pos, d.N, amp / d.hs.L[0], amp / d.hs.L[1], amp / d.hs.L[2]);
        MDL_HIP(hipGetLastError());
        k_quantize<<<(unsigned)((d.N + BLOCK - 1) / BLOCK), BLOCK, 0, d.stream>>>(d.a, d.N);
        MDL_HIP(hipGetLastError());
    }
    double best = 1e300;
    int best_lpa = 16;
    int const cand[] = {4, 8, 16, 32, 64};
    int const saved = d.lpa;
    for (int lpa : cand) {
        d.lpa = lpa;
        d.gridA = d.gridB = d.grid_pairs(lpa);
        d.launch_forces(true);   // warm-up
        MDL_HIP(hipEventRecord(e0, d.stream));
        for (int r = 0; r < 3; ++r)
            d.launch_forces(true);
        MDL_HIP(hipEventRecord(e1, d.stream));
        MDL_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        MDL_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_lpa = lpa;
        }
    }
    (void)saved;
    MDL_HIP(hipMemcpyAsync(d.a.pos, saved_pos, sizeof(double4) * (size_t)d.N,
                           hipMemcpyDeviceToDevice, d.stream));
    k_quantize<<<(unsigned)((d.N + BLOCK - 1) / BLOCK), BLOCK, 0, d.stream>>>(d.a, d.N);
    MDL_HIP(hipGetLastError());
    {
        int const zero = 0;   // a displaced density above rho_max is no error of the run
        MDL_HIP(hipMemcpyAsync(&d.gs->err, &zero, sizeof(int), hipMemcpyHostToDevice, d.stream));
        MDL_HIP(hipStreamSynchronize(d.stream));
    }
    MDL_HIP(hipFree(saved_pos));
    d.lpa = best_lpa;
    d.gridA = d.gridB = d.grid_pairs(best_lpa);
    {
        std::lock_guard<std::mutex> lock(g_lpa_mutex);
        g_lpa_cache[key] = best_lpa;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return MDL_OK;
}

static mdl_status create(const mdl_config *c, Driver &d) {
    if (!c || !c->positions || !c->velocities || !c->types || !c->mass || !c->pair)
        return bad("mdl_driver_create: null argument");
    if (c->N < 1 || c->N > (int64_t)1 << 30)
        return bad("mdl_driver_create: N out of range");
    if (c->ntypes < 1 || c->ntypes > MAX_TYPES) {
        set_error("mdl_driver_create: %d species; the GPU path supports 1..%d", c->ntypes,
                  MAX_TYPES);
        return MDL_ERR_UNSUPPORTED;
    }
    if (c->th_M < 1 || c->th_M > MAX_CHAIN || c->bth_M < 1 || c->bth_M > MAX_CHAIN || !c->th_q ||
        !c->th_p_eta || !c->bth_q || !c->bth_p_eta) {
        set_error("mdl_driver_create: chain lengths must be 1..%d", MAX_CHAIN);
        return MDL_ERR_UNSUPPORTED;
    }
    if (!(c->dt > 0.))
        return bad("mdl_driver_create: dt must be > 0");
    for (int a = 0; a < 3; ++a)
        if (!(c->box[a] > 0.))
            return bad("mdl_driver_create: box lengths must be > 0");
    d.N = (int)c->N;
    d.ntypes = c->ntypes;
    d.npair = c->ntypes * (c->ntypes + 1) / 2;
    d.has_eam = c->has_eam ? 1 : 0;
    d.precision = c->precision;
    if (d.precision != MDL_FP64 && d.precision != MDL_MIXED)
        return bad("mdl_driver_create: precision must be MDL_FP64 or MDL_MIXED");
    if (d.precision == MDL_MIXED && !d.has_eam) {
        set_error("mdl_driver_create: the mixed-precision path needs EAM tables");
        return MDL_ERR_UNSUPPORTED;
    }
    if (d.has_eam && (!c->rho_tab || !c->F_tab || !c->eam_rcut || !c->eam_rhomax))
        return bad("mdl_driver_create: EAM tables missing");
    // md_driver.cpp:91-106: positions inside the box, types in range
    for (int64_t i = 0; i < c->N; ++i) {
        const double *p = c->positions + 3 * i;
        if (std::fabs(p[0]) >= 0.5 * c->box[0] || std::fabs(p[1]) >= 0.5 * c->box[1] ||
            std::fabs(p[2]) >= 0.5 * c->box[2]) {
            set_error("bad initial position");
            return MDL_ERR_POSITION;
        }
        if (c->types[i] >= (uint32_t)c->ntypes) {
            set_error("bad type");
            return MDL_ERR_TYPE;
        }
    }
    if (c->ensemble != MDL_NVE && c->ensemble != MDL_NVT && c->ensemble != MDL_NPT) {
        set_error("bad ensamble %d", c->ensemble);
        return MDL_ERR_ENSEMBLE;
    }
    // tables: one interpolation kind, a common r grid and a common rho grid
    std::vector<const HostTable *> rtabs;
    for (int p = 0; p < d.npair; ++p)
        rtabs.push_back(reinterpret_cast<const HostTable *>(c->pair[p]));
    if (d.has_eam)
        for (int t = 0; t < d.ntypes; ++t)
            rtabs.push_back(reinterpret_cast<const HostTable *>(c->rho_tab[t]));
    for (auto *t : rtabs)
        if (!t)
            return bad("mdl_driver_create: null table");
    d.kind = rtabs[0]->kind;
    if (analytic_kind(d.kind) && d.has_eam) {
        set_error("analytic pair potentials (Morse, Lennard-Jones) run without EAM tables "
                  "on the GPU path");
        return MDL_ERR_UNSUPPORTED;
    }
    for (auto *t : rtabs)
        if (t->kind != d.kind || t->h != rtabs[0]->h || t->n != rtabs[0]->n) {
            set_error("the GPU path needs all r-tables on one grid with one interpolation");
            return MDL_ERR_UNSUPPORTED;
        }
    d.rc_pair = c->pair_rcut[0];
    d.rmin = c->pair_rmin ? c->pair_rmin[0] : 0.;
    for (int p = 1; p < d.npair; ++p)
        if (c->pair_rcut[p] != d.rc_pair || (c->pair_rmin && c->pair_rmin[p] != d.rmin)) {
            set_error("the GPU path needs one cut-off and one r_min for all pair tables");
            return MDL_ERR_UNSUPPORTED;
        }
    d.rc2 = c->pair_rcut[0] * c->pair_rcut[0];   // md_driver.cpp:42 (pot[0]->rcut)
    d.rc_rho = 1e300;
    double rhomax[MAX_TYPES] = {0};
    if (d.has_eam) {
        d.rc_rho = c->eam_rcut[0];
        const HostTable *f0 = reinterpret_cast<const HostTable *>(c->F_tab[0]);
        for (int t = 0; t < d.ntypes; ++t) {
            const HostTable *ftab = reinterpret_cast<const HostTable *>(c->F_tab[t]);
            if (!ftab)
                return bad("mdl_driver_create: null F table");
            if (c->eam_rcut[t] != d.rc_rho || ftab->kind != d.kind || ftab->h != f0->h ||
                ftab->n != f0->n) {
                set_error("the GPU path needs one rho cut-off and one F(rho) grid for all species");
                return MDL_ERR_UNSUPPORTED;
            }
            rhomax[t] = c->eam_rhomax[t];
        }
    }
    // interleaved r records: [interval][function][S]; analytic forms hold one
    // record of PAIR_PARAMS parameters per pair slot
    int const S = rtabs[0]->stride;
    int const nf = (int)rtabs.size();
    int64_t const nr = analytic_kind(d.kind) ? 1 : rtabs[0]->n;
    std::vector<double> rrec((size_t)nr * nf * S);
    for (int64_t i = 0; i < nr; ++i)
        for (int f = 0; f < nf; ++f)
            memcpy(&rrec[((size_t)i * nf + f) * S], &rtabs[f]->rec[(size_t)i * S],
                   sizeof(double) * S);
    MDL_HIP(hipGetDevice(&d.device));
    MDL_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    if (mdl_status s = d.alloc(&d.d_rrec, rrec.size()))
        return s;
    MDL_HIP(hipMemcpy(d.d_rrec, rrec.data(), sizeof(double) * rrec.size(), hipMemcpyHostToDevice));
    d.rt.rec = d.d_rrec;
    d.rt.nf = nf;
    d.rt.n = (int)nr;
    d.rt.inv_h = rtabs[0]->inv_h;
    d.rt.inv_h2 = rtabs[0]->inv_h2;
    if (d.has_eam) {
        const HostTable *f0 = reinterpret_cast<const HostTable *>(c->F_tab[0]);
        std::vector<double> frec((size_t)f0->n * d.ntypes * S);
        for (int64_t i = 0; i < f0->n; ++i)
            for (int t = 0; t < d.ntypes; ++t)
                memcpy(&frec[((size_t)i * d.ntypes + t) * S],
                       &reinterpret_cast<const HostTable *>(c->F_tab[t])->rec[(size_t)i * S],
                       sizeof(double) * S);
        if (mdl_status s = d.alloc(&d.d_frec, frec.size()))
            return s;
        MDL_HIP(hipMemcpy(d.d_frec, frec.data(), sizeof(double) * frec.size(),
                          hipMemcpyHostToDevice));
        d.ft.rec = d.d_frec;
        d.ft.n = (int)f0->n;
        d.ft.inv_h = f0->inv_h;
        d.ft.inv_h2 = f0->inv_h2;
        for (int t = 0; t < d.ntypes; ++t)
            d.ft.rhomax[t] = rhomax[t];
    }

    if (d.precision == MDL_MIXED) {
        // FP32 records in eV / Angstrom units: value scale per function, and for
        // B-spline records the independent f' and f'' cubics also carry the
        // argument's unit change (cm -> A for r, cm^-3 -> A^-3 for rho)
        double const EV = 1.6021773e-12;
        int const SF = d.kind == MDL_QUINTIC ? 8 : 12;
        std::vector<float> rf((size_t)nr * nf * SF, 0.f);
        for (int64_t i = 0; i < nr; ++i)
            for (int f = 0; f < nf; ++f) {
                double const sv = f < d.npair ? 1. / EV : 1e-24;
                const double *src = &rtabs[f]->rec[(size_t)i * S];
                float *dst = &rf[((size_t)i * nf + f) * SF];
                if (d.kind == MDL_QUINTIC)
                    for (int c = 0; c < 6; ++c)
                        dst[c] = (float)(src[c] * sv);
                else
                    for (int c = 0; c < 12; ++c)
                        dst[c] = (float)(src[c] * sv * (c < 4 ? 1. : (c < 8 ? 1e-8 : 1e-16)));
            }
        if (mdl_status st = d.alloc(&d.d_rrec_f, rf.size()))
            return st;
        MDL_HIP(hipMemcpy(d.d_rrec_f, rf.data(), sizeof(float) * rf.size(), hipMemcpyHostToDevice));
        // the force pass's fused records [interval][pair slot][PairRecF::NF]:
        // phi of the slot, then the derivative parts of rho of its two species,
        // the same floats as the records above
        {
            int const NFP = d.kind == MDL_QUINTIC ? 16 : 32;
            std::vector<float> pf((size_t)nr * d.npair * NFP, 0.f);
            for (int64_t i = 0; i < nr; ++i)
                for (int b = 0; b < d.ntypes; ++b)
                    for (int a = 0; a <= b; ++a) {
                        int const slot = b * (b + 1) / 2 + a;
                        const float *phi = &rf[((size_t)i * nf + slot) * SF];
                        const float *ra = &rf[((size_t)i * nf + d.npair + a) * SF];
                        const float *rb = &rf[((size_t)i * nf + d.npair + b) * SF];
                        float *dst = &pf[((size_t)i * d.npair + slot) * NFP];
                        if (d.kind == MDL_QUINTIC) {
                            for (int k = 0; k < 6; ++k)
                                dst[k] = phi[k];
                            for (int k = 0; k < 5; ++k) {
                                dst[6 + k] = ra[1 + k];
                                dst[11 + k] = rb[1 + k];
                            }
                        } else {
                            for (int k = 0; k < 12; ++k)
                                dst[k] = phi[k];
                            for (int k = 0; k < 8; ++k) {
                                dst[12 + k] = ra[4 + k];
                                dst[20 + k] = rb[4 + k];
                            }
                        }
                    }
            if (mdl_status st = d.alloc(&d.d_prec_f, pf.size()))
                return st;
            MDL_HIP(hipMemcpy(d.d_prec_f, pf.data(), sizeof(float) * pf.size(),
                              hipMemcpyHostToDevice));
        }
        d.rtf.rec = d.d_rrec_f;
        d.rtf.prec = d.d_prec_f;
        d.rtf.nf = nf;
        d.rtf.n = (int)nr;
        d.rtf.inv_h = (float)(rtabs[0]->inv_h * 1e-8);
        d.rtf.inv_h2 = (float)(rtabs[0]->inv_h2 * 1e-16);
        const HostTable *f0 = reinterpret_cast<const HostTable *>(c->F_tab[0]);
        std::vector<float> ff((size_t)f0->n * d.ntypes * SF, 0.f);
        for (int64_t i = 0; i < f0->n; ++i)
            for (int t = 0; t < d.ntypes; ++t) {
                const double *src =
                    &reinterpret_cast<const HostTable *>(c->F_tab[t])->rec[(size_t)i * S];
                float *dst = &ff[((size_t)i * d.ntypes + t) * SF];
                if (d.kind == MDL_QUINTIC)
                    for (int k = 0; k < 6; ++k)
                        dst[k] = (float)(src[k] / EV);
                else
                    for (int k = 0; k < 12; ++k)
                        dst[k] = (float)(src[k] / EV * (k < 4 ? 1. : (k < 8 ? 1e24 : 1e48)));
            }
        if (mdl_status st = d.alloc(&d.d_frec_f, ff.size()))
            return st;
        MDL_HIP(hipMemcpy(d.d_frec_f, ff.data(), sizeof(float) * ff.size(), hipMemcpyHostToDevice));
        d.ftf.rec = d.d_frec_f;
        d.ftf.n = (int)f0->n;
        d.ftf.inv_h = f0->inv_h * 1e24;
        d.ftf.inv_hf = (float)(f0->inv_h * 1e24);
        d.ftf.inv_h2f = (float)(f0->inv_h2 * 1e48);
        for (int t = 0; t < d.ntypes; ++t)
            d.ftf.rhomax[t] = rhomax[t] * 1e-24;
    }

    // neighbour list geometry: the list is exact at every step (rebuilt on
    // request and whenever a pair outside it could have entered r_c)
    double const rc = c->pair_rcut[0];
    double skin = c->neighbors_list ? c->r_skin : 1.2;
    if (const char *env = getenv("MDL_SKIN"))
        skin = atof(env);
    skin = std::min(std::max(skin, 1.02), 2.0);
    double const r_list = rc * skin;
    int nc[3];
    bool cells = true;
    for (int a = 0; a < 3; ++a) {
        nc[a] = (int)std::floor(c->box[a] / (r_list * 1.02));
        if (nc[a] < 3)
            cells = false;
    }
    if (getenv("MDL_BRUTE_LIST"))
        cells = false;
    d.use_cells = cells;
    double const V = c->box[0] * (c->box[1] * c->box[2]);
    double const expect = 4.18879020478639 * r_list * r_list * r_list * (double)d.N / V;
    int stride = (int)std::ceil(1.5 * expect + 32.);
    stride = std::min(stride, std::max(d.N - 1, 1));
    stride = (stride + 7) / 8 * 8;
    d.l.stride = stride;
    d.l.ncells = cells ? nc[0] * nc[1] * nc[2] : 0;

    int const N = d.N;
    AtomArrays &a = d.a;
#define ALLOC(p, n)                                                                            \
    if (mdl_status s_ = d.alloc(&(p), (n)))                                                    \
        return s_;
    ALLOC(a.pos, N);
    ALLOC(a.qpos, N);
    ALLOC(a.vx, N);
    ALLOC(a.vy, N);
    ALLOC(a.vz, N);
    ALLOC(a.fx, N);
    ALLOC(a.fy, N);
    ALLOC(a.fz, N);
    ALLOC(a.type, N);
    ALLOC(a.cref, N);
    ALLOC(a.mref, N);
    ALLOC(a.fprime, N);
    ALLOC(a.geam, (size_t)4 * N);
    ALLOC(a.fprime_f, N);
    ALLOC(a.Femb, N);
    ALLOC(a.rho, N);
    ALLOC(a.epot, N);
    ALLOC(d.l.nbr, (size_t)N * stride);
    ALLOC(d.l.nnbr, N);
    ALLOC(d.l.cnbr, (size_t)N * stride);
    ALLOC(d.l.ccnt, N);
    ALLOC(d.l.mnbr, (size_t)N * stride);
    ALLOC(d.l.mcnt, N);
    MDL_HIP(hipMemset(d.l.mcnt, 0, sizeof(int) * N));
    if (cells) {
        ALLOC(d.l.cell_of, N);
        ALLOC(d.l.cell_count, d.l.ncells);
        ALLOC(d.l.cell_start, d.l.ncells + 1);
        ALLOC(d.l.cell_fill, d.l.ncells);
        ALLOC(d.l.cell_atoms, N);
        MDL_HIP(hipMemset(d.l.cell_count, 0, sizeof(int) * d.l.ncells));
        MDL_HIP(hipMemset(d.l.cell_fill, 0, sizeof(int) * d.l.ncells));
    }
    ALLOC(d.pA, (size_t)PART_STRIDE * PA_N);
    ALLOC(d.pB, (size_t)PART_STRIDE * PB_N);
    ALLOC(d.pK, (size_t)PART_STRIDE * PK_N);
    ALLOC(d.d_stage, (size_t)3 * N);
    ALLOC(d.d_totals, NTOTAL);
    ALLOC(d.d_bar, 4);
    ALLOC(d.d_ticket_fin, 4);
    MDL_HIP(hipMemset(d.d_bar, 0, 4 * sizeof(unsigned)));
    MDL_HIP(hipMemset(d.d_ticket_fin, 0, 4 * sizeof(unsigned)));
    static_assert(GRID_ATOM_MAX <= DMAX_MID, "displacement maxima overlap");
    ALLOC(d.d_dmax, 2 * DMAX_MID);
    MDL_HIP(hipMemset(d.d_dmax, 0, sizeof(double) * 2 * DMAX_MID));
    ALLOC(d.gs, 1);
#undef ALLOC
    MDL_HIP(hipMemset(a.fx, 0, sizeof(double) * N));
    MDL_HIP(hipMemset(a.fy, 0, sizeof(double) * N));
    MDL_HIP(hipMemset(a.fz, 0, sizeof(double) * N));
    MDL_HIP(hipMemset(a.Femb, 0, sizeof(double) * N));
    MDL_HIP(hipMemset(a.fprime, 0, sizeof(double2) * N));
    MDL_HIP(hipMemset(a.geam, 0, sizeof(double2) * 4 * N));
    MDL_HIP(hipMemset(a.rho, 0, sizeof(double) * N));

    // scaled coordinates (md_driver.cpp:162-173) and SoA velocities
    std::vector<double4> pos(N);
    std::vector<double> vx(N), vy(N), vz(N);
    std::vector<int> typ(N);
    auto nint = [](double x) { return (double)(x < 0. ? (int)(x - 0.5) : (int)(x + 0.5)); };
    for (int i = 0; i < N; ++i) {
        double s[3];
        for (int k = 0; k < 3; ++k) {
            double const u = c->positions[3 * i + k] / c->box[k];
            s[k] = u - nint(u);
        }
        typ[i] = (int)c->types[i];
        pos[i] = make_double4(s[0], s[1], s[2], (double)typ[i]);
        vx[i] = c->velocities[3 * i];
        vy[i] = c->velocities[3 * i + 1];
        vz[i] = c->velocities[3 * i + 2];
    }
    d.init_pos.assign(c->positions, c->positions + 3 * (size_t)N);
    d.type_count.assign((size_t)c->ntypes, 0);
    for (int i = 0; i < N; ++i)
        ++d.type_count[(size_t)typ[i]];
    MDL_HIP(hipMemcpy(a.pos, pos.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
    MDL_HIP(hipMemcpy(a.cref, pos.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
    MDL_HIP(hipMemcpy(a.mref, pos.data(), sizeof(double4) * N, hipMemcpyHostToDevice));
    MDL_HIP(hipMemcpy(a.vx, vx.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    MDL_HIP(hipMemcpy(a.vy, vy.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    MDL_HIP(hipMemcpy(a.vz, vz.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    MDL_HIP(hipMemcpy(a.type, typ.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    k_quantize<<<(unsigned)((N + BLOCK - 1) / BLOCK), BLOCK, 0, d.stream>>>(a, N);
    MDL_HIP(hipGetLastError());

    DevState &s = d.hs;
    memset(&s, 0, sizeof s);
    s.N = N;
    s.ntypes = d.ntypes;
    s.ensemble = c->ensemble;
    s.coupling = c->coupling;
    s.has_eam = d.has_eam;
    s.sample = 1;
    s.Nf = c->Nf;
    s.alpha = 1. + 3. / c->Nf;
    s.dt = c->dt;
    s.fatal = c->fatal_conservation;
    for (int t = 0; t < d.ntypes; ++t)
        s.mass[t] = c->mass[t];
    for (int k = 0; k < 3; ++k)
        s.L[k] = s.L0[k] = s.L_drift[k] = s.L_build[k] = s.drift_L[k] = c->box[k];
    s.th.T = c->th_T;
    s.th.Nf = c->th_Nf;
    s.th.M = c->th_M;
    for (int j = 0; j < c->th_M; ++j) {
        s.th.q[j] = c->th_q[j];
        s.th.iq[j] = 1.0 / c->th_q[j];
        s.th.p_eta[j] = c->th_p_eta[j];
    }
    s.bth.T = c->bth_T;
    s.bth.Nf = c->bth_Nf;
    s.bth.M = c->bth_M;
    for (int j = 0; j < c->bth_M; ++j) {
        s.bth.q[j] = c->bth_q[j];
        s.bth.iq[j] = 1.0 / c->bth_q[j];
        s.bth.p_eta[j] = c->bth_p_eta[j];
    }
    s.bar_P = c->bar_P;
    s.bar_W = c->bar_W;
    s.peps = c->bar_p_eps;
    s.vscale = 1.;
    s.kick_s = 1.;
    s.mixed = d.precision == MDL_MIXED ? 1 : 0;
    s.unit_E = s.mixed ? 1.6021773e-12 : 1.;
    s.unit_L = s.mixed ? 1.6021773e-12 * 1e16 : 1.;
    s.r_list = r_list;
    s.rc = rc;
    s.skin = r_list - rc;
    // mid list: 0.3 A beyond r_c (half the skin when that is thinner); it is
    // refiltered when an atom may have moved 0.15 A since the last filtering
    // mixed: pass B scans the mid list itself, so its shell is kept thin
    s.r_mid = rc + std::min(d.precision == MDL_MIXED ? 0.15e-8 : 0.3e-8, 0.5 * (r_list - rc));
    s.mid_build = 1;
    for (int k = 0; k < 3; ++k)
        s.L_mid[k] = c->box[k];
    s.force_rebuild = 0;
    for (int k = 0; k < 3; ++k)
        s.ncell[k] = cells ? nc[k] : 1;
    MDL_HIP(hipMemcpy(d.gs, &s, sizeof(DevState), hipMemcpyHostToDevice));
    // the uploads and memsets above ran on the null stream, which the driver's
    // non-blocking stream does not wait for
    MDL_HIP(hipDeviceSynchronize());

    d.lpa = c->lanes_per_atom;
    if (d.lpa != 0 && d.lpa != 4 && d.lpa != 8 && d.lpa != 16 && d.lpa != 32 && d.lpa != 64)
        return bad("lanes_per_atom must be 0 (autotune), 4, 8, 16, 32 or 64");
    bool const tune = d.lpa == 0;
    if (tune)
        d.lpa = 16;
    d.gridA = d.gridB = d.grid_pairs(d.lpa);
    d.gridK = std::min((N + BLOCK - 1) / BLOCK, GRID_ATOM_MAX);
    {   // k_list's grid barriers need every block resident at once: at most
        // what the device holds of it (occupancy x CUs), and at most 512
        // blocks, since each barrier costs a serialised atomic per block
        int per_cu = 0, cus = 0;
        MDL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_list, BLOCK, 0));
        MDL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d.device));
        if (per_cu < 1 || cus < 1) {
            set_error("k_list does not fit on the device");
            return MDL_ERR_DEVICE;
        }
        int64_t const want = d.use_cells
                                 ? std::max<int64_t>(d.l.ncells, (N + BLOCK - 1) / BLOCK)
                                 : ((int64_t)N * LIST_LANES + BLOCK - 1) / BLOCK;
        int64_t const cap = std::min<int64_t>(512, (int64_t)per_cu * cus);
        d.gridL = (int)std::max<int64_t>(1, std::min<int64_t>(want, cap));
        d.gridL_solo = d.gridL;
        d.list_cap = (int64_t)per_cu * cus;
    }

    // constructor: list, forces, thermodynamics (md_driver.cpp:112-118)
    d.launch_build();
    d.build_pending = false;
    MDL_HIP(hipGetLastError());
    if (mdl_status st = d.sync_state())
        return st;
    if (mdl_status st = d.check_errors())
        return st;
    if (tune && d.N >= 2048)
        if (mdl_status st = autotune(d))
            return st;
    d.launch_forces(true);
    k_kick2<true><<<d.gridK, BLOCK, 0, d.stream>>>(d.a, N, d.gs, 1, d.has_eam, d.pK);
    d.launch_finish(0, nullptr, -1);
    MDL_HIP(hipGetLastError());
    if (mdl_status st = d.sync_state())
        return st;
    return d.check_errors();
}

}  // namespace mdl

using namespace mdl;

// Mixed path: atom i's pairs within r_c, counted over its mid list with the
// passes' own decision (pair_r2 of the fixed-point pair vector against rc2f).
__global__ __launch_bounds__(256) void k_count_within_rc(AtomArrays a, ListArrays l, int N,
                                                        const DevState *__restrict__ gs,
                                                        float rc2f, double *__restrict__ out) {
    int const i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N)
        return;
    float const S[3] = {(float)(gs->L[0] * 1e8) * 0x1p-32f, (float)(gs->L[1] * 1e8) * 0x1p-32f,
                        (float)(gs->L[2] * 1e8) * 0x1p-32f};
    int4 const qi = a.qpos[i];
    const int *nb = l.mnbr + (size_t)i * l.stride;
    int const m = l.mcnt[i];
    int n = 0;
    for (int t = 0; t < m; ++t) {
        float x, y, z;
        qmin_image(qi, a.qpos[nb[t]], S, x, y, z);
        n += pair_r2(x, y, z) > rc2f ? 0 : 1;
    }
    out[i] = (double)n;
}

// The pair list a control-variate kernel reads: 0 the pairs pass A kept within
// the potential's cut-off (FP64 path), 1 the full list (before any force
// evaluation, or for a cut-off beyond the potential's), 2 the mid list (mixed
// path, whose passes take the r_c decision themselves). Each holds every pair
// within the potential's cut-off, in list order; the readers test r.
__device__ __forceinline__ int pair_list(const ListArrays &l, int which, int i, const int *&nb) {
    nb = (which == 1 ? l.nbr : which == 2 ? l.mnbr : l.cnbr) + (size_t)i * l.stride;
    return which == 1 ? l.nnbr[i] : which == 2 ? l.mcnt[i] : l.ccnt[i];
}

// Control-variate descriptors of the current configuration from the pairs of
// list `which` (pair_list) within the basis cut-off.
// One wave per two atoms, lane k%32 = radial function k; lengths in A.
template <bool AXES>
__global__ __launch_bounds__(256) void k_cv_desc_list(AtomArrays a, ListArrays l, int N,
                                                      const DevState *__restrict__ gs,
                                                      double to_A, CvBasis basis, int which,
                                                      float *__restrict__ D,
                                                      float *__restrict__ J,
                                                      float *__restrict__ Lp,
                                                      float *__restrict__ La) {
    int const lane = threadIdx.x & 63, k = lane & 31;
    int const i = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 2 + (lane >> 5);
    bool const active = i < N;
    int const ia = active ? i : 0;
    // pair vectors from the fixed-point positions (A per unit: L 2^-32 to_A)
    float const S[3] = {(float)(gs->L[0] * to_A) * 0x1p-32f, (float)(gs->L[1] * to_A) * 0x1p-32f,
                        (float)(gs->L[2] * to_A) * 0x1p-32f};
    int4 const ci = a.qpos[ia];
    const int *nb;
    int const cnt0 = pair_list(l, which, ia, nb);
    int const cnt = active ? cnt0 : 0;
    CvAcc<AXES> acc;
    // the 32 lanes of an atom gather 32 pairs at once (one load round trip per
    // 32 pairs), then take them one by one from the lane that holds each, in
    // list order
    int const half = lane & 32;
    for (int q0 = 0; q0 < cnt; q0 += 32) {
        int const q = q0 + k;
        float x = 0.f, y = 0.f, z = 0.f, r = 1.f;
        int sp = 0, ok = 0;
        if (q < cnt) {
            int4 const cj = a.qpos[nb[q]];
            qmin_image(ci, cj, S, x, y, z);
            float const r2 = x * x + y * y + z * z;
            ok = r2 < basis.rc2 ? 1 : 0;
            r = ok ? __builtin_amdgcn_sqrtf(r2) : 1.f;
            sp = cj.w;
        }
        int const m = min(32, cnt - q0);
        for (int t = 0; t < m; ++t) {
            int const src = half + t;
            if (__shfl(ok, src, 64))
                acc.add(basis, k, __shfl(sp, src, 64), __shfl(x, src, 64), __shfl(y, src, 64),
                        __shfl(z, src, 64), __shfl(r, src, 64));
        }
    }
    if (active)
        acc.store(i, k, N, D, J, Lp, La);
}

// Atom i's candidate pairs of the step just launched, list `which` (pair_list:
// a superset of its pairs within the cut-off, in list order; the readers test
// r): count[i] = their number and nbr[i][0..min(count, width)); a count beyond
// width tells the caller that the row was cut. One wave per atom.
__global__ __launch_bounds__(256) void k_cv_capture_pairs(ListArrays l, int N, int which,
                                                         int width, int *__restrict__ nbr,
                                                         int *__restrict__ count) {
    int const lane = threadIdx.x & 63;
    int const i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= N)
        return;
    const int *src;
    int const cnt = pair_list(l, which, i, src);
    int *dst = nbr + (size_t)i * width;
    int const m = min(cnt, width);
    for (int t = lane; t < m; t += 64)
        dst[t] = src[t];
    if (lane == 0)
        count[i] = cnt;
}

// The control-variate inputs of the step just launched, into the next frame
// slot (descriptors from the compacted list, positions and forces, and the pair
// lists when asked for), on the driver's stream.
static void cv_capture(Driver &d) {
    Driver::CvCapture &c = d.cvc;
    int const j = c.next % c.slots;
    ++c.next;
    size_t const nf = (size_t)d.N * 2 * CV_NB;
    int const waves = (d.N + 1) / 2;
    k_cv_desc_list<false><<<(waves * 64 + 255) / 256, 256, 0, d.stream>>>(
        d.a, d.l, d.N, d.gs, 1e8, c.basis, c.list, c.D + j * nf, c.J + j * 3 * nf,
        c.L + j * nf, nullptr);
    unsigned const g = (d.N + 255) / 256;
    k_materialize<<<g, 256, 0, d.stream>>>(d.a, d.N, d.gs, MDL_A_POSITIONS,
                                           c.pos + (size_t)j * 3 * d.N);
    k_materialize<<<g, 256, 0, d.stream>>>(d.a, d.N, d.gs, MDL_A_FORCES,
                                           c.force + (size_t)j * 3 * d.N);
    if (c.nbr)
        k_cv_capture_pairs<<<(d.N + 3) / 4, 256, 0, d.stream>>>(
            d.l, d.N, c.list, c.width, c.nbr + (size_t)j * d.N * c.width,
            c.ncnt + (size_t)j * d.N);
}

// ---------------------------------------------------------------------------
// Pair correlation (md_driver.cpp:732-863): every pair i < j whose minimum-image
// distance r is below rmaxg counts in bin floor(r / drg), in the total and in
// its unordered species pair. The distance is formed as the reference forms it
// (no contraction, correctly rounded sqrt and division), so the integer counts
// are the reference's. Block (I, J), J >= I, takes atoms [I T, I T + T) against
// [J T, J T + T); counts go to LDS when they fit, then to 64-bit global counters.
// ---------------------------------------------------------------------------
constexpr int GR_T = 256;

__device__ __forceinline__ double pair_distance_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return sqrt(dx * dx + dy * dy + dz * dz);
}

__global__ __launch_bounds__(GR_T) void k_pair_hist(const double4 *__restrict__ pos, int N,
                                                    const DevState *__restrict__ gs, int nbin,
                                                    double drg, double rmaxg, int want_total,
                                                    int want_partial, int ntypes,
                                                    unsigned long long *__restrict__ cnt,
                                                    int use_lds) {
    extern __shared__ unsigned hist[];   // [1 + npair][nbin] when use_lds
    __shared__ double4 pj[GR_T];
    int const I = blockIdx.x, J = blockIdx.y;
    if (J < I)
        return;
    int const npair = ntypes * (ntypes + 1) / 2;
    int const nh = (1 + npair) * nbin;
    if (use_lds)
        for (int q = threadIdx.x; q < nh; q += GR_T)
            hist[q] = 0u;
    int const j0 = J * GR_T;
    if (j0 + (int)threadIdx.x < N)
        pj[threadIdx.x] = pos[j0 + threadIdx.x];
    __syncthreads();
    double const L[3] = {gs->L[0], gs->L[1], gs->L[2]};
    int const i = I * GR_T + threadIdx.x;
    if (i < N) {
        double4 const ci = pos[i];
        int const ti = (int)ci.w;
        int const jn = min(GR_T, N - j0);
        for (int u = 0; u < jn; ++u) {
            if (j0 + u <= i)
                continue;
            double4 const cj = pj[u];
            double dx, dy, dz;
            min_image(ci, cj, L, dx, dy, dz);
            double const r = pair_distance_exact(dx, dy, dz);
            if (!(r < rmaxg))
                continue;
            int const k = (int)floor(r / drg);
            if (k >= nbin)
                continue;
            int const tj = (int)cj.w;
            int const tmax = max(ti, tj), tmin = min(ti, tj);
            int const slot = 1 + ((tmax * (tmax + 1)) >> 1) + tmin;
            if (use_lds) {
                if (want_total)
                    atomicAdd(&hist[k], 1u);
                if (want_partial)
                    atomicAdd(&hist[slot * nbin + k], 1u);
            } else {
                if (want_total)
                    atomicAdd(&cnt[k], 1ull);
                if (want_partial)
                    atomicAdd(&cnt[(size_t)slot * nbin + k], 1ull);
            }
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int q = threadIdx.x; q < nh; q += GR_T)
            if (hist[q])
                atomicAdd(&cnt[q], (unsigned long long)hist[q]);
    }
}

// The exponential operator on n points: form 0 expx_operator (barostat,
// chains longer than the register forms), 1 the chain sweeps' wide form
// (expx_any), 2 their series form (op_series, |x| <= EXPX_SMALL).
__global__ __launch_bounds__(BLOCK) void k_expx_eval(int form, int64_t n,
                                                    const double *__restrict__ x,
                                                    double *__restrict__ expx,
                                                    double *__restrict__ op) {
    int64_t const i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n)
        return;
    double e, o;
    if (form == 0) {
        expx_operator(x[i], e, o);
    } else if (form == 1) {
        expx_any(x[i], e, o);
    } else {
        o = op_series(x[i]);
        e = fma(-x[i], o, 1.);
    }
    expx[i] = e;
    op[i] = o;
}

__global__ __launch_bounds__(256) void k_int_to_double(const int *__restrict__ in, int n,
                                                       double *__restrict__ out) {
    int const i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        out[i] = (double)in[i];
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long *__restrict__ p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
        p[k] = 0ull;
}

// One g(r) call: its bins, rows (row 0 the total, row 1 + q species pair q)
// and the normalisation constants of md_driver.cpp:749-751 and :786-809.
struct GrCall {
    int nbin = 0, nrow = 0, ntypes = 0;
    bool want_total = false, want_partial = false;
    double drg = 0., rmaxg = 0., V = 0., Np = 0.;
    double pref[MAX_TYPES * (MAX_TYPES + 1) / 2] = {};   // V / (4 pi n_pairs drg), 0 if none
    size_t nh = 0;
};

static mdl_status gr_prepare(Driver *d, const double *rg, int32_t nbin, int32_t mode,
                             double *gr, double *gr_partial, GrCall &c, size_t max_nh = 0) {
    if (!d || !rg || nbin < 2 || mode < 0 || mode > 2)
        return bad("mdl_driver_pair_correlation: bad arguments");
    c.want_total = mode != 2;
    c.want_partial = mode != 0;
    if ((c.want_total && !gr) || (c.want_partial && !gr_partial))
        return bad("mdl_driver_pair_correlation: missing output array");
    c.drg = rg[1] - rg[0];
    if (!(c.drg > 0.))
        return bad("mdl_driver_pair_correlation: rg must increase");
    c.rmaxg = rg[nbin - 1] + 0.5 * c.drg;   // md_driver.cpp:41
    MDL_HIP(hipSetDevice(d->device));
    if (d->failed)
        return bad("the driver stopped on an earlier error");
    c.nbin = nbin;
    c.ntypes = d->ntypes;
    int const npair = d->ntypes * (d->ntypes + 1) / 2;
    c.nrow = 1 + npair;
    c.nh = (size_t)c.nrow * nbin;
    if (max_nh && c.nh > max_nh) {
        set_error("mdl_driver_pair_correlation_noalloc: %zu bin counters ((1 + species "
                  "pairs) x nbin) exceed the scratch of %zu; use mdl_driver_pair_correlation",
                  c.nh, max_nh);
        return MDL_ERR_ARG;
    }
    double L[3] = {0., 0., 0.};
    MDL_HIP(hipMemcpyAsync(L, d->gs->L, sizeof L, hipMemcpyDeviceToHost, d->stream));
    MDL_HIP(hipStreamSynchronize(d->stream));
    c.V = L[0] * (L[1] * L[2]);
    c.Np = (double)d->N;
    int q = 0;
    for (int b = 0; b < d->ntypes; ++b)
        for (int a = 0; a <= b; ++a, ++q) {
            double const Na = (double)d->type_count[a], Nb = (double)d->type_count[b];
            double pc = 0.;
            if (a == b)
                pc = Na > 1 ? Na * (Na - 1.) / 2. : 0.;
            else
                pc = Na * Nb;
            c.pref[q] = pc > 0. ? c.V / (4. * M_PI * pc * c.drg) : 0.;
        }
    return MDL_OK;
}

// Zero the counters (a fill kernel: the call may follow captured work) and
// count every pair i < j into them.
static hipError_t gr_count(Driver *d, const GrCall &c, unsigned long long *cnt) {
    unsigned const fill = (unsigned)std::min<size_t>((c.nh + 255) / 256, 1024);
    k_fill_u64<<<fill, 256, 0, d->stream>>>(cnt, c.nh);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    size_t const lds = c.nh * sizeof(unsigned);
    int const use_lds = lds <= 48 * 1024 ? 1 : 0;
    int const T = (d->N + GR_T - 1) / GR_T;
    k_pair_hist<<<dim3(T, T), GR_T, use_lds ? lds : 0, d->stream>>>(
        d->a.pos, d->N, d->gs, c.nbin, c.drg, c.rmaxg, c.want_total ? 1 : 0,
        c.want_partial ? 1 : 0, d->ntypes, cnt, use_lds);
    return hipGetLastError();
}

// Where row `row` goes (nullptr: the mode does not compute it).
static double *gr_row_out(const GrCall &c, int row, double *gr, double *gr_partial) {
    if (row == 0)
        return c.want_total ? gr : nullptr;
    return c.want_partial ? gr_partial + (size_t)(row - 1) * c.nbin : nullptr;
}

// Bins k0 .. k0 + n - 1 of one row from their counts cnt[0 .. n-1], in the
// reference's order of operations.
static void gr_normalize(const GrCall &c, int row, int k0, int n,
                         const unsigned long long *cnt, const double *rg, double *out) {
    for (int u = 0; u < n; ++u) {
        int const k = k0 + u;
        if (row == 0)
            out[k] = (double)cnt[u] * (c.V / (2. * M_PI * c.Np * c.Np * rg[k] * rg[k] * c.drg));
        else if (c.pref[row - 1] > 0.)
            out[k] = (double)cnt[u] * (c.pref[row - 1] / (rg[k] * rg[k]));
        else
            out[k] = 0.;
    }
}

// Per-device scratch of mdl_driver_pair_correlation_noalloc: the bin counters
// live in this file-scope array (one copy per GPU, held until the process
// ends), so the call needs no device allocation. Its device address is taken
// once per device with hipGetSymbolAddress; host code uses that address, never
// the array's name. All calls on one GPU share it: they are serialised by
// g_gr_scratch_mutex, and calls from other code that overlap would race.
static __device__ __align__(256) unsigned long long scratch[MDL_GR_SCRATCH_COUNTS];
static std::mutex g_gr_scratch_mutex;
static std::map<int, unsigned long long *> g_gr_scratch_ptr;

static mdl_status gr_scratch(int device, unsigned long long **out) {
    auto it = g_gr_scratch_ptr.find(device);
    if (it != g_gr_scratch_ptr.end()) {
        *out = it->second;
        return MDL_OK;
    }
    void *ptr = nullptr;
    hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                        "hipGetSymbolAddress(scratch)");
    g_gr_scratch_ptr[device] = static_cast<unsigned long long *>(ptr);
    *out = static_cast<unsigned long long *>(ptr);
    return MDL_OK;
}

// ---------------------------------------------------------------------------
// One captured graph for a chunk of steps of several drivers: each driver's
// launches form a branch of its own (its stream joins the capture through a
// fork event), so one launch advances every replica by graph_steps steps.
// Cached per (drivers, stride, build interval, chunk length); a driver's
// destruction or a change of its chunk length drops the entries with it.
// ---------------------------------------------------------------------------
struct JointGraph {
    std::vector<uint64_t> ids;
    int64_t stride = 0, rebuild = 0, steps = 0;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr, after = nullptr;
    std::vector<hipEvent_t> joins;
    std::vector<hipEvent_t> ready;   // each driver's stream before a launch
    std::vector<double *> rec;       // each driver's records of one chunk
    std::vector<int64_t> nrec;

    void release() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        for (hipEvent_t e : joins)
            if (e)
                (void)hipEventDestroy(e);
        for (hipEvent_t e : ready)
            if (e)
                (void)hipEventDestroy(e);
        ready.clear();
        if (fork)
            (void)hipEventDestroy(fork);
        if (after)
            (void)hipEventDestroy(after);
        exec = nullptr;
        fork = after = nullptr;
        joins.clear();
    }
};

std::mutex g_joint_mutex;
std::vector<JointGraph> g_joint;
uint64_t g_next_uid = 1;

void joint_forget(uint64_t uid) {
    std::lock_guard<std::mutex> lock(g_joint_mutex);
    for (size_t k = 0; k < g_joint.size();) {
        JointGraph &j = g_joint[k];
        if (std::find(j.ids.begin(), j.ids.end(), uid) != j.ids.end()) {
            j.release();
            g_joint.erase(g_joint.begin() + (long)k);
        } else {
            ++k;
        }
    }
}

static Driver *D(mdl_driver *p) { return reinterpret_cast<Driver *>(p); }

// k_list's in-kernel build waits at grid barriers until all of its blocks are
// resident. Drivers run side by side on one device may all be inside that
// build at once, so their k_list grids together must fit the device: with
// `share` drivers on it, each takes at most 1/share of what it holds. The
// list a build makes does not depend on the grid (each cell or atom row is
// built whole by one block), so the results do not change. Graphs captured
// with the old grid are dropped.
static void share_list_grid(Driver *d, int share) {
    int64_t const g = std::max<int64_t>(
        1, std::min<int64_t>(d->gridL_solo, d->list_cap / std::max(share, 1)));
    if (g != d->gridL) {
        d->gridL = (int)g;
        d->drop_graph();
        joint_forget(d->uid);
    }
}

extern "C" {

mdl_status mdl_driver_create(const mdl_config *cfg, mdl_driver **out) {
    if (!out)
        return bad("mdl_driver_create: null output");
    *out = nullptr;
    auto d = std::make_unique<Driver>();
    mdl_status s = create(cfg, *d);
    if (s != MDL_OK)
        return s;
    {
        std::lock_guard<std::mutex> lock(g_joint_mutex);
        d->uid = g_next_uid++;
    }
    *out = reinterpret_cast<mdl_driver *>(d.release());
    return MDL_OK;
}

mdl_status mdl_driver_destroy(mdl_driver *p) {
    Driver *d = D(p);
    if (!d)
        return MDL_OK;
    (void)hipSetDevice(d->device);
    (void)hipStreamSynchronize(d->stream);
    joint_forget(d->uid);
    delete d;
    return MDL_OK;
}

static mdl_status guard(Driver *d) {
    if (!d)
        return bad("null driver");
    if (d->failed) {
        set_error("the driver failed earlier and accepts no more steps");
        return MDL_ERR_STATE;
    }
    MDL_HIP(hipSetDevice(d->device));
    return MDL_OK;
}

mdl_status mdl_driver_integrate(mdl_driver *p, int64_t nsteps) {
    Driver *d = D(p);
    if (mdl_status s = guard(d))
        return s;
    share_list_grid(d, 1);
    int const smp = d->sample != 0 ? 1 : 0;
    if (nsteps > 0)
        k_pre<<<1, BLOCK, 0, d->stream>>>(d->gs, smp);
    for (int64_t k = 0; k < nsteps; ++k) {
        bool const build = d->build_pending;
        d->build_pending = false;
        MDL_HIP(d->launch_step(smp != 0, nullptr, nullptr, build, k + 1 < nsteps ? smp : -1));
    }
    if (mdl_status s = d->sync_state())
        return s;
    return d->check_errors();
}

}  // extern "C"

// One run() of a driver, split into begin / one launch per step / end so that
// the launches of several drivers (each on its own stream) can be interleaved
// from one host thread.
struct RunPlan {
    Driver *d = nullptr;
    int64_t nsteps = 0, stride = 1, rebuild = 0, k = 0, irec = 0, t0 = 0;
    int ntime = 0;
    std::vector<char> timed;   // timing slots whose events this run recorded
    double *records = nullptr, *drec = nullptr;
    int64_t *n_records = nullptr;

    bool sampled(int64_t cyc) const { return (cyc + 1) % stride == 0; }
    bool forced(int64_t cyc) const { return rebuild > 0 && cyc % rebuild == 0; }

    mdl_status begin() {
        if (mdl_status s = guard(d))
            return s;
        if (stride < 1)
            return bad("sample_stride must be >= 1");
        int64_t nsamp = 0;
        for (int64_t j = 0; j < nsteps; ++j)
            if ((d->cycle_host + j + 1) % stride == 0)
                ++nsamp;
        if (records && nsamp > 0) {
            if (nsamp > d->run_rec_cap) {
                if (mdl_status st = d->alloc(&d->run_rec, (size_t)(MDL_NRECORD * nsamp)))
                    return st;
                d->run_rec_cap = nsamp;
            }
            drec = d->run_rec;
        }
        // events around the passes of up to 64 steps
        ntime = (int)std::min<int64_t>(nsteps, 64);
        while ((int)d->ev.size() < 4 * ntime) {
            hipEvent_t e;
            MDL_HIP(hipEventCreate(&e));
            d->ev.push_back(e);
        }
        t0 = nsteps - ntime;   // time the last steps (after any warm-up)
        timed.assign((size_t)ntime, 0);
        if (nsteps > 0)
            k_pre<<<1, BLOCK, 0, d->stream>>>(d->gs, sampled(d->cycle_host) ? 1 : 0);
        MDL_HIP(hipGetLastError());
        return MDL_OK;
    }

    // graph_steps steps from here as one graph replay: the cycle is aligned, no
    // build is pending, the build interval and stride fit the chunk, and the
    // final step of the run stays out (it ends without the next first half)
    bool graph_fits() const {
        int64_t const G = d->graph_steps;
        return G > 0 && k + G < nsteps && d->cycle_host % G == 0 && !d->build_pending &&
               G % stride == 0 && (rebuild <= 0 || G % rebuild == 0) && d->cvc.slots == 0;
    }

    mdl_status graph_step() {
        if (!d->gexec || d->g_stride != stride || d->g_rebuild != rebuild)
            if (mdl_status s = d->capture_graph(stride, rebuild))
                return s;
        MDL_HIP(hipGraphLaunch(d->gexec, d->stream));
        if (drec && d->g_nrec > 0)
            MDL_HIP(hipMemcpyAsync(drec + MDL_NRECORD * irec, d->g_rec,
                                   sizeof(double) * MDL_NRECORD * d->g_nrec,
                                   hipMemcpyDeviceToDevice, d->stream));
        irec += d->g_nrec;
        k += d->graph_steps;
        d->cycle_host += d->graph_steps;
        return MDL_OK;
    }

    // Launch step k (asynchronous), or a captured chunk of steps when allowed.
    mdl_status step(bool graphs = true) {
        if (graphs && graph_fits())
            return graph_step();
        int64_t const cyc = d->cycle_host;
        bool const s = sampled(cyc);
        double *ro = (drec && s) ? drec + MDL_NRECORD * irec : nullptr;
        if (s)
            ++irec;
        hipEvent_t *tev = nullptr;
        if (k >= t0) {
            tev = &d->ev[4 * (k - t0)];
            (void)hipEventRecord(tev[3], d->stream);
            timed[(size_t)(k - t0)] = 1;
        }
        int const ns = k + 1 < nsteps ? (sampled(cyc + 1) ? 1 : 0) : -1;
        bool const build = d->build_pending || forced(cyc);
        d->build_pending = false;
        ++k;
        MDL_HIP(d->launch_step(s, ro, tev, build, ns));
        if (s && d->cvc.slots > 0) {
            cv_capture(*d);
            MDL_HIP(hipGetLastError());
        }
        return MDL_OK;
    }

    mdl_status end() {
        if (drec) {
            hipError_t e = hipMemcpyAsync(records, drec, sizeof(double) * MDL_NRECORD * irec,
                                          hipMemcpyDeviceToHost, d->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(d->stream);
            drec = nullptr;   // the driver's buffer, reused by the next run
            if (e != hipSuccess)
                return hip_fail(e, "mdl_driver_run records");
        }
        if (n_records)
            *n_records = irec;
        if (mdl_status s = d->sync_state())
            return s;
        // pass timings: [3] step start, [0] before A, [1] after A, [2] after B
        double ta = 0, tb = 0, tsum = 0;
        int nt = 0;
        // steps replayed from a graph recorded no events: only recorded pairs
        int nstep = 0;
        for (int j = 0; j < ntime; ++j) {
            if (!timed[(size_t)j])
                continue;
            hipEvent_t *e = &d->ev[4 * j];
            float x = 0, y = 0;
            if (hipEventElapsedTime(&x, e[0], e[1]) == hipSuccess &&
                hipEventElapsedTime(&y, e[1], e[2]) == hipSuccess) {
                ta += x;
                tb += y;
                ++nt;
            }
            if (j + 1 < ntime && timed[(size_t)j + 1]) {
                float z = 0;
                if (hipEventElapsedTime(&z, e[3], d->ev[4 * (j + 1) + 3]) == hipSuccess) {
                    tsum += z;
                    ++nstep;
                }
            }
        }
        (void)hipGetLastError();   // a failed elapsed-time query must not linger
        if (nt > 0) {
            d->t_dens = ta / nt;
            d->t_force = tb / nt;
        }
        if (nstep > 0)
            d->t_step = tsum / nstep;
        return d->check_errors();
    }
};

// Every plan can take the same graph chunk now (same chunk length, device,
// position in the run, and each fits).
static bool joint_fits(std::vector<RunPlan> &plans) {
    RunPlan const &p0 = plans[0];
    for (RunPlan const &r : plans)
        if (!r.graph_fits() || r.k != p0.k || r.d->graph_steps != p0.d->graph_steps ||
            r.d->device != p0.d->device)
            return false;
    return true;
}

static mdl_status joint_capture(std::vector<RunPlan> &plans, JointGraph &j) {
    size_t const n = plans.size();
    Driver *d0 = plans[0].d;
    j.stride = plans[0].stride;
    j.rebuild = plans[0].rebuild;
    j.steps = d0->graph_steps;
    j.rec.assign(n, nullptr);
    j.nrec.assign(n, 0);
    j.joins.assign(n, nullptr);
    j.ready.assign(n, nullptr);
    for (size_t i = 0; i < n; ++i)
        MDL_HIP(hipEventCreateWithFlags(&j.ready[i], hipEventDisableTiming));
    MDL_HIP(hipEventCreateWithFlags(&j.fork, hipEventDisableTiming));
    MDL_HIP(hipEventCreateWithFlags(&j.after, hipEventDisableTiming));
    for (size_t i = 0; i < n; ++i) {
        j.ids.push_back(plans[i].d->uid);
        MDL_HIP(hipEventCreateWithFlags(&j.joins[i], hipEventDisableTiming));
        if (mdl_status st = plans[i].d->chunk_records(j.stride, j.nrec[i]))
            return st;
        j.rec[i] = plans[i].d->g_rec;
    }
    hipGraph_t graph = nullptr;
    MDL_HIP(hipStreamBeginCapture(d0->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipEventRecord(j.fork, d0->stream);
    for (size_t i = 1; i < n && e == hipSuccess; ++i)
        e = hipStreamWaitEvent(plans[i].d->stream, j.fork, 0);
    for (size_t i = 0; i < n && e == hipSuccess; ++i)
        e = plans[i].d->chunk_launches(j.stride, j.rebuild);
    for (size_t i = 1; i < n && e == hipSuccess; ++i) {
        e = hipEventRecord(j.joins[i], plans[i].d->stream);
        if (e == hipSuccess)
            e = hipStreamWaitEvent(d0->stream, j.joins[i], 0);
    }
    hipError_t const e2 = hipStreamEndCapture(d0->stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (graph)
            (void)hipGraphDestroy(graph);
        return hip_fail(e != hipSuccess ? e : e2, "joint step graph capture");
    }
    e = hipGraphInstantiate(&j.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        j.exec = nullptr;
        return hip_fail(e, "joint step graph instantiation");
    }
    return MDL_OK;
}

// One launch of the joint chunk; afterwards every driver's stream waits for it.
static mdl_status joint_step(std::vector<RunPlan> &plans) {
    size_t const n = plans.size();
    Driver *d0 = plans[0].d;
    MDL_HIP(hipSetDevice(d0->device));
    std::lock_guard<std::mutex> lock(g_joint_mutex);
    JointGraph *j = nullptr;
    for (JointGraph &c : g_joint) {
        if (c.ids.size() != n || c.stride != plans[0].stride || c.rebuild != plans[0].rebuild ||
            c.steps != d0->graph_steps)
            continue;
        bool same = true;
        for (size_t i = 0; i < n && same; ++i)
            same = c.ids[i] == plans[i].d->uid;
        if (same) {
            j = &c;
            break;
        }
    }
    if (!j) {
        g_joint.emplace_back();
        j = &g_joint.back();
        if (mdl_status st = joint_capture(plans, *j)) {
            j->release();
            g_joint.pop_back();
            return st;
        }
    }
    // the graph starts on the first driver's stream: it must also come after
    // the work already queued on every other driver's stream (their previous
    // steps), which the fork inside the graph does not wait for
    for (size_t i = 1; i < n; ++i) {
        MDL_HIP(hipEventRecord(j->ready[i], plans[i].d->stream));
        MDL_HIP(hipStreamWaitEvent(d0->stream, j->ready[i], 0));
    }
    MDL_HIP(hipGraphLaunch(j->exec, d0->stream));
    MDL_HIP(hipEventRecord(j->after, d0->stream));
    for (size_t i = 0; i < n; ++i) {
        RunPlan &r = plans[i];
        if (i > 0)
            MDL_HIP(hipStreamWaitEvent(r.d->stream, j->after, 0));
        if (r.drec && j->nrec[i] > 0)
            MDL_HIP(hipMemcpyAsync(r.drec + MDL_NRECORD * r.irec, j->rec[i],
                                   sizeof(double) * MDL_NRECORD * j->nrec[i],
                                   hipMemcpyDeviceToDevice, r.d->stream));
        r.irec += j->nrec[i];
        r.k += j->steps;
        r.d->cycle_host += j->steps;
    }
    return MDL_OK;
}

// ---------------------------------------------------------------------------
// Several systems per launch. One table entry per replica holds what its
// kernels take; a block map gives every block of a launch its replica and its
// block index within that replica's share, which is the replica's own grid
// (the same blocks and partial sums as its own launches, so the results are
// bit for bit those of the replica run alone with the same lanes per atom).
// The list check runs as one block per replica.
// ---------------------------------------------------------------------------
struct RepArgs {
    AtomArrays a;
    ListArrays l;
    RTablesF rt;
    FTablesF ft;
    DevState *gs;
    double *pA, *pB, *pK, *totals, *dmax, *rec;
    unsigned *ticket, *bar;
    int N, npair, ntypes, gridA, gridB, gridK, use_cells, has_eam;
    double rc2A;
    float rmin;
};

__global__ __launch_bounds__(BLOCK) void k_kick_drift_rep(const RepArgs *__restrict__ R,
                                                         const int2 *__restrict__ map) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    kick_drift_body(r.a, r.N, r.gs, r.dmax, Blk{m.y, r.gridK});
}

__global__ __launch_bounds__(BLOCK) void k_list_rep(const RepArgs *__restrict__ R) {
    RepArgs const &r = R[blockIdx.x];
    list_body(r.a, r.l, r.N, r.gs, r.bar, r.dmax, r.gridK, r.use_cells, Blk{0, 1});
}

template <int KIND, int LPA, bool SAMPLE>
__global__ __launch_bounds__(BLOCK) void k_density_rep(const RepArgs *__restrict__ R,
                                                      const int2 *__restrict__ map) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    density_mx_body<KIND, LPA, SAMPLE>(r.a, r.l, r.rt, r.ft, r.N, r.npair, r.ntypes, r.rc2A,
                                       r.gs, r.pA, Blk{m.y, r.gridA});
}

template <int KIND, int LPA, bool SAMPLE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void k_force_rep(const RepArgs *__restrict__ R, const int2 *__restrict__ map) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    force_mx_body<KIND, LPA, SAMPLE>(r.a, r.l, r.rt, r.N, r.npair, r.rmin, (float)r.rc2A, r.gs, r.pB,
                                     Blk{m.y, r.gridB});
}

// `init` and `mode` stay kernel arguments as in k_kick2 and k_finish: a
// literal would let the compiler fold the branches they select and contract
// the remaining sums differently, and the kinetic sums would then differ from
// the replica's own launches in the last bit.
template <bool SAMPLE>
__global__ __launch_bounds__(BLOCK) void k_kick2_rep(const RepArgs *__restrict__ R,
                                                    const int2 *__restrict__ map, int init) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    kick2_body<SAMPLE>(r.a, r.N, r.gs, init, r.has_eam, r.pK, Blk{m.y, r.gridK});
}

// The end of the step in two launches: the totals (one block per component
// and replica, blockIdx.y the replica), then the scalar tail (one block per
// replica), the same sums and tail as k_finish.
__global__ __launch_bounds__(BLOCK) void k_finish_sum_rep(const RepArgs *__restrict__ R) {
    RepArgs const &r = R[blockIdx.y];
    if (r.gs->err)
        return;
    double const t =
        finish_sum(r.gs, r.pA, r.gridA, r.pB, r.gridB, r.pK, r.gridK, (int)blockIdx.x);
    if (threadIdx.x == 0)
        r.totals[blockIdx.x] = t;
}

// irec < 0: not a sampled step; else the record index, counted from *base
// when base is given (the records of a replayed chunk of steps)
__global__ __launch_bounds__(BLOCK) void k_finish_tail_rep(const RepArgs *__restrict__ R, int mode,
                                                          int irec, const int *__restrict__ base,
                                                          int next_sample) {
    RepArgs const &r = R[blockIdx.x];
    if (r.gs->err)
        return;
    int64_t const ix = irec < 0 ? -1 : (int64_t)irec + (base ? *base : 0);
    double *ro = (ix >= 0 && r.rec) ? r.rec + (size_t)MDL_NRECORD * ix : nullptr;
    finish_tail(r.gs, r.totals, nullptr, mode, ro, next_sample);
}

__global__ void k_set_int(int *p, int v) {
    if (threadIdx.x == 0)
        *p = v;
}

// A forced list build of every replica, in five launches (those of
// Driver::launch_build, each over the blocks of all replicas). Replicas with
// a brute-force list take no part in the scan, scatter and sort launches; the
// maps give each launch its blocks.
__global__ __launch_bounds__(BLOCK) void k_rb_bin_rep(const RepArgs *__restrict__ R,
                                                     const int2 *__restrict__ map,
                                                     const int *__restrict__ nblk) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    if (r.gs->err)
        return;
    if (m.y == 0 && threadIdx.x == 0) {
        rb_begin(r.gs, r.use_cells);
        r.gs->nlist_entries = 0;
    }
    if (r.use_cells) {
        ListArrays l = r.l;
        rb_bin(r.a, l, r.N, r.gs, m.y * BLOCK, nblk[m.x] * BLOCK);
    }
}

__global__ __launch_bounds__(BLOCK) void k_rb_scan_rep(const RepArgs *__restrict__ R,
                                                      const int *__restrict__ which) {
    RepArgs const &r = R[which[blockIdx.x]];
    if (r.gs->err)
        return;
    __shared__ int sc[BLOCK];
    ListArrays l = r.l;
    rb_scan(l, sc);
}

__global__ __launch_bounds__(BLOCK) void k_rb_scatter_rep(const RepArgs *__restrict__ R,
                                                         const int2 *__restrict__ map,
                                                         const int *__restrict__ nblk) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    if (r.gs->err)
        return;
    ListArrays l = r.l;
    rb_scatter(l, r.N, m.y * BLOCK, nblk[m.x] * BLOCK);
}

__global__ __launch_bounds__(BLOCK) void k_rb_sort_rep(const RepArgs *__restrict__ R,
                                                      const int2 *__restrict__ map) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    if (r.gs->err)
        return;
    __shared__ int v[CELL_MAX];
    ListArrays l = r.l;
    rb_sort_cell(l, m.y, v, r.gs);
}

__global__ __launch_bounds__(BLOCK) void k_rb_build_rep(const RepArgs *__restrict__ R,
                                                       const int2 *__restrict__ map,
                                                       const int *__restrict__ nblk) {
    int2 const m = map[blockIdx.x];
    RepArgs const &r = R[m.x];
    if (r.gs->err)
        return;
    __shared__ ListSmem sm;
    if (threadIdx.x == 0)
        sm.entries = 0ull;
    double const L[3] = {r.gs->L[0], r.gs->L[1], r.gs->L[2]};
    double const r_list = r.gs->r_list;
    AtomArrays a = r.a;
    ListArrays l = r.l;
    if (r.use_cells)
        rb_build_cell(a, l, m.y, r.gs, L, r_list, sm);
    else
        rb_build_brute(a, l, r.N, r.gs, L, r_list, (m.y * BLOCK + threadIdx.x) / LIST_LANES,
                       nblk[m.x] * BLOCK / LIST_LANES, &sm.entries);
    rb_flush_entries(r.gs, sm.entries);
}

using RepPassFn = void (*)(const RepArgs *, const int2 *);

template <int KIND, bool S> static RepPassFn rep_density_fn(int lpa) {
    switch (lpa) {
    case 4: return k_density_rep<KIND, 4, S>;
    case 8: return k_density_rep<KIND, 8, S>;
    case 16: return k_density_rep<KIND, 16, S>;
    case 32: return k_density_rep<KIND, 32, S>;
    default: return k_density_rep<KIND, 64, S>;
    }
}
template <int KIND, bool S> static RepPassFn rep_force_fn(int lpa) {
    switch (lpa) {
    case 4: return k_force_rep<KIND, 4, S>;
    case 8: return k_force_rep<KIND, 8, S>;
    case 16: return k_force_rep<KIND, 16, S>;
    case 32: return k_force_rep<KIND, 32, S>;
    default: return k_force_rep<KIND, 64, S>;
    }
}
static RepPassFn pick_density_rep(int kind, int lpa, bool s) {
    if (kind == MDL_QUINTIC)
        return s ? rep_density_fn<MDL_QUINTIC, true>(lpa) : rep_density_fn<MDL_QUINTIC, false>(lpa);
    return s ? rep_density_fn<MDL_BSPLINE, true>(lpa) : rep_density_fn<MDL_BSPLINE, false>(lpa);
}
static RepPassFn pick_force_rep(int kind, int lpa, bool s) {
    if (kind == MDL_QUINTIC)
        return s ? rep_force_fn<MDL_QUINTIC, true>(lpa) : rep_force_fn<MDL_QUINTIC, false>(lpa);
    return s ? rep_force_fn<MDL_BSPLINE, true>(lpa) : rep_force_fn<MDL_BSPLINE, false>(lpa);
}

// The drivers can share launches (checked before any of them is touched).
static mdl_status replicas_check(const std::vector<Driver *> &ds) {
    Driver *d0 = ds[0];
    for (Driver *d : ds) {
        if (mdl_status s = guard(d))
            return s;
        if (d->precision != MDL_MIXED)
            return bad("mdl_replicas_run: every driver must use mixed precision");
        if (d->device != d0->device)
            return bad("mdl_replicas_run: the drivers are on different devices");
        if (d->cvc.slots > 0)
            return bad("mdl_replicas_run: a driver captures control-variate inputs");
        if (d->gridK > 4 * BLOCK)
            return bad("mdl_replicas_run: a replica is too large for the one-block list check");
        if (d->cycle_host != d0->cycle_host)
            return bad("mdl_replicas_run: the drivers are not at the same cycle");
    }
    return MDL_OK;
}

static mdl_status replicas_run(std::vector<RunPlan> &plans, int64_t nsteps) {
    size_t const n = plans.size();
    Driver *d0 = plans[0].d;
    // a requested build is taken by the replicas' own launches first
    bool pending = false;
    for (RunPlan &r : plans)
        pending = pending || r.d->build_pending;
    if (pending)
        for (RunPlan &r : plans)
            if (r.k < nsteps)
                if (mdl_status st = r.step(false))
                    return st;
    for (RunPlan &r : plans)
        MDL_HIP(hipStreamSynchronize(r.d->stream));
    hipStream_t const st = d0->stream;
    std::vector<RepArgs> tab(n);
    std::vector<int2> mapK, mapBin, mapScat, mapSort, mapBuild;
    std::vector<int> nblkN(n), nblkB(n), scanWhich;
    // the pair passes: one launch per group of replicas with the same
    // interpolation kind and lanes per atom
    struct Group {
        int kind, lpa;
        std::vector<int2> mapA, mapB;
        size_t atA = 0, atB = 0;
    };
    std::vector<Group> groups;
    for (size_t i = 0; i < n; ++i) {
        Driver *d = plans[i].d;
        RepArgs &e = tab[i];
        e.a = d->a;
        e.l = d->l;
        e.rt = d->rtf;
        e.ft = d->ftf;
        e.gs = d->gs;
        e.pA = d->pA;
        e.pB = d->pB;
        e.pK = d->pK;
        e.totals = d->d_totals;
        e.dmax = d->d_dmax;
        e.rec = plans[i].drec;
        e.ticket = d->d_ticket_fin;
        e.bar = d->d_bar;
        e.N = d->N;
        e.npair = d->npair;
        e.ntypes = d->ntypes;
        e.gridA = d->gridA;
        e.gridB = d->gridB;
        e.gridK = d->gridK;
        e.use_cells = d->use_cells ? 1 : 0;
        e.has_eam = d->has_eam;
        e.rc2A = d->rc2 * 1e16;
        e.rmin = (float)(d->rmin * 1e8);
        Group *g = nullptr;
        for (Group &x : groups)
            if (x.kind == d->kind && x.lpa == d->lpa)
                g = &x;
        if (!g) {
            groups.push_back(Group{d->kind, d->lpa, {}, {}});
            g = &groups.back();
        }
        for (int b = 0; b < d->gridA; ++b)
            g->mapA.push_back(make_int2((int)i, b));
        for (int b = 0; b < d->gridB; ++b)
            g->mapB.push_back(make_int2((int)i, b));
        for (int b = 0; b < d->gridK; ++b)
            mapK.push_back(make_int2((int)i, b));
        // the forced build's launches (Driver::launch_build's grids)
        int const gN = (d->N + BLOCK - 1) / BLOCK;
        int const gB = d->use_cells ? d->l.ncells
                                    : (int)(((int64_t)d->N * LIST_LANES + BLOCK - 1) / BLOCK);
        nblkN[i] = gN;
        nblkB[i] = gB;
        for (int b = 0; b < (d->use_cells ? gN : 1); ++b)
            mapBin.push_back(make_int2((int)i, b));
        if (d->use_cells) {
            scanWhich.push_back((int)i);
            for (int b = 0; b < gN; ++b)
                mapScat.push_back(make_int2((int)i, b));
            for (int c = 0; c < d->l.ncells; ++c)
                mapSort.push_back(make_int2((int)i, c));
        }
        for (int b = 0; b < gB; ++b)
            mapBuild.push_back(make_int2((int)i, b));
    }
    // the tables in the lead driver's buffer: the entries, the int2 maps, the
    // int arrays, then the record base of a replayed chunk
    std::vector<int2> all;
    std::vector<size_t> at;
    for (const std::vector<int2> *v : {&mapK, &mapBin, &mapScat, &mapSort, &mapBuild}) {
        at.push_back(all.size());
        all.insert(all.end(), v->begin(), v->end());
    }
    for (Group &g : groups) {
        g.atA = all.size();
        all.insert(all.end(), g.mapA.begin(), g.mapA.end());
        g.atB = all.size();
        all.insert(all.end(), g.mapB.begin(), g.mapB.end());
    }
    std::vector<int> ints(nblkN);
    ints.insert(ints.end(), nblkB.begin(), nblkB.end());
    ints.insert(ints.end(), scanWhich.begin(), scanWhich.end());
    size_t const oMap = (sizeof(RepArgs) * n + 255) & ~(size_t)255;
    size_t const oInt = oMap + sizeof(int2) * all.size();
    size_t const oBase = (oInt + sizeof(int) * ints.size() + 7) & ~(size_t)7;
    size_t const bytes = oBase + sizeof(int64_t);
    std::vector<char> image(bytes, 0);
    std::memcpy(image.data(), tab.data(), sizeof(RepArgs) * n);
    std::memcpy(image.data() + oMap, all.data(), sizeof(int2) * all.size());
    std::memcpy(image.data() + oInt, ints.data(), sizeof(int) * ints.size());
    std::vector<int64_t> shape = {(int64_t)n,           (int64_t)mapK.size(),
                                  (int64_t)mapBin.size(), (int64_t)mapScat.size(),
                                  (int64_t)mapSort.size(), (int64_t)mapBuild.size(),
                                  (int64_t)scanWhich.size()};
    for (const Group &g : groups)
        shape.insert(shape.end(), {g.kind, g.lpa, (int64_t)g.atA, (int64_t)g.mapA.size(),
                                   (int64_t)g.atB, (int64_t)g.mapB.size()});
    RepSet &rs = d0->rep;
    if (bytes > rs.cap) {
        rs.release();
        if (hipMalloc(&rs.buf, bytes) != hipSuccess) {
            rs.buf = nullptr;
            return bad("mdl_replicas_run: out of device memory");
        }
        rs.cap = bytes;
    }
    if (rs.image != image || rs.shape != shape) {
        rs.drop_graph();   // its launches were sized for the old tables
        MDL_HIP(hipMemcpy(rs.buf, image.data(), bytes, hipMemcpyHostToDevice));
        rs.image.swap(image);
        rs.shape.swap(shape);
    }
    char *const buf = static_cast<char *>(rs.buf);
    RepArgs const *dtab = reinterpret_cast<RepArgs *>(buf);
    int2 const *dmap = reinterpret_cast<int2 *>(buf + oMap);
    int const *dint = reinterpret_cast<int *>(buf + oInt);
    int *dbase = reinterpret_cast<int *>(buf + oBase);
    int2 const *mK = dmap + at[0], *mBin = dmap + at[1], *mScat = dmap + at[2],
               *mSort = dmap + at[3], *mBuild = dmap + at[4];
    int const *dN = dint, *dB = dint + n, *dWhich = dint + 2 * n;

    // one batched step; irec as k_finish_tail_rep takes it
    auto launch = [&](bool smp, bool build, int irec, const int *base, int ns) {
        k_kick_drift_rep<<<(unsigned)mapK.size(), BLOCK, 0, st>>>(dtab, mK);
        if (build) {
            k_rb_bin_rep<<<(unsigned)mapBin.size(), BLOCK, 0, st>>>(dtab, mBin, dN);
            if (!scanWhich.empty()) {
                k_rb_scan_rep<<<(unsigned)scanWhich.size(), BLOCK, 0, st>>>(dtab, dWhich);
                k_rb_scatter_rep<<<(unsigned)mapScat.size(), BLOCK, 0, st>>>(dtab, mScat, dN);
                k_rb_sort_rep<<<(unsigned)mapSort.size(), BLOCK, 0, st>>>(dtab, mSort);
            }
            k_rb_build_rep<<<(unsigned)mapBuild.size(), BLOCK, 0, st>>>(dtab, mBuild, dB);
        } else
            k_list_rep<<<(unsigned)n, BLOCK, 0, st>>>(dtab);
        for (const Group &g : groups)
            pick_density_rep(g.kind, g.lpa, smp)<<<(unsigned)g.mapA.size(), BLOCK, 0, st>>>(
                dtab, dmap + g.atA);
        for (const Group &g : groups)
            pick_force_rep(g.kind, g.lpa, smp)<<<(unsigned)g.mapB.size(), BLOCK, 0, st>>>(
                dtab, dmap + g.atB);
        if (smp)
            k_kick2_rep<true><<<(unsigned)mapK.size(), BLOCK, 0, st>>>(dtab, mK, 0);
        else
            k_kick2_rep<false><<<(unsigned)mapK.size(), BLOCK, 0, st>>>(dtab, mK, 0);
        k_finish_sum_rep<<<dim3(NTOTAL, (unsigned)n), BLOCK, 0, st>>>(dtab);
        k_finish_tail_rep<<<(unsigned)n, BLOCK, 0, st>>>(dtab, 1, smp ? irec : -1, base, ns);
        return hipGetLastError();
    };

    // graph_steps (the lead driver's) batched steps as one replay: from an
    // aligned cycle whose chunk the stride and build interval divide, never
    // the run's last step; the chunk's records go from the record base on
    RunPlan &p0 = plans[0];
    int64_t const G = d0->graph_steps;
    auto chunk_fits = [&]() {
        return G > 0 && p0.k + G < nsteps && d0->cycle_host % G == 0 && G % p0.stride == 0 &&
               (p0.rebuild <= 0 || G % p0.rebuild == 0);
    };
    auto capture = [&]() -> mdl_status {
        rs.drop_graph();
        int64_t const c0 = d0->cycle_host;
        int nrec = 0;
        hipGraph_t graph = nullptr;
        MDL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int64_t j = 0; j < G && e == hipSuccess; ++j) {
            int64_t const cyc = c0 + j;
            bool const smp = p0.sampled(cyc);
            e = launch(smp, p0.forced(cyc), smp ? nrec++ : -1, dbase, p0.sampled(cyc + 1) ? 1 : 0);
        }
        hipError_t const e2 = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph)
                (void)hipGraphDestroy(graph);
            return hip_fail(e != hipSuccess ? e : e2, "batched step graph capture");
        }
        e = hipGraphInstantiate(&rs.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            rs.exec = nullptr;
            return hip_fail(e, "batched step graph instantiation");
        }
        rs.stride = p0.stride;
        rs.rebuild = p0.rebuild;
        rs.steps = G;
        rs.nrec = nrec;
        return MDL_OK;
    };

    hipError_t e = hipSuccess;
    while (e == hipSuccess && p0.k < nsteps) {
        int64_t const cyc = d0->cycle_host;
        if (chunk_fits()) {
            if (!rs.exec || rs.stride != p0.stride || rs.rebuild != p0.rebuild || rs.steps != G)
                if (mdl_status s = capture())
                    return s;
            k_set_int<<<1, 64, 0, st>>>(dbase, (int)p0.irec);
            e = hipGetLastError();
            if (e == hipSuccess)
                e = hipGraphLaunch(rs.exec, st);
            for (RunPlan &r : plans) {
                r.d->cycle_host += G;
                r.k += G;
                r.irec += rs.nrec;
            }
            continue;
        }
        bool const smp = p0.sampled(cyc);
        int const ns = p0.k + 1 < nsteps ? (p0.sampled(cyc + 1) ? 1 : 0) : -1;
        e = launch(smp, p0.forced(cyc), (int)p0.irec, nullptr, ns);
        for (RunPlan &r : plans) {
            ++r.d->cycle_host;
            ++r.k;
            if (smp)
                ++r.irec;
        }
    }
    hipError_t const e2 = hipStreamSynchronize(st);
    if (e != hipSuccess || e2 != hipSuccess)
        return hip_fail(e != hipSuccess ? e : e2, "mdl_replicas_run");
    return MDL_OK;
}

extern "C" {

mdl_status mdl_driver_run(mdl_driver *p, int64_t nsteps, int64_t sample_stride,
                          int64_t rebuild_every, double *records, int64_t *n_records) {
    RunPlan r;
    r.d = D(p);
    r.nsteps = nsteps;
    r.stride = sample_stride;
    r.rebuild = rebuild_every;
    r.records = records;
    r.n_records = n_records;
    if (r.d)
        share_list_grid(r.d, 1);
    if (mdl_status s = r.begin())
        return s;
    while (r.k < nsteps)
        if (mdl_status s = r.step())
            return s;
    return r.end();
}

mdl_status mdl_drivers_run(mdl_driver *const *drivers, int32_t n, int64_t nsteps,
                           int64_t sample_stride, int64_t rebuild_every, double *const *records,
                           int64_t *n_records) {
    if (!drivers || n < 1)
        return bad("no drivers");
    std::vector<RunPlan> plans((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t j = 0; j < i; ++j)
            if (drivers[j] == drivers[i])
                return bad("a driver appears twice in mdl_drivers_run");
        // every driver is checked before any run begins (begin() advances the state)
        if (mdl_status s = guard(D(drivers[i])))
            return s;
    }
    if (sample_stride < 1)
        return bad("sample_stride must be >= 1");
    std::map<int, int> on_device;
    for (int32_t i = 0; i < n; ++i)
        ++on_device[D(drivers[i])->device];
    for (int32_t i = 0; i < n; ++i)
        share_list_grid(D(drivers[i]), on_device[D(drivers[i])->device]);
    for (int32_t i = 0; i < n; ++i) {
        RunPlan &r = plans[(size_t)i];
        r.d = D(drivers[i]);
        r.nsteps = nsteps;
        r.stride = sample_stride;
        r.rebuild = rebuild_every;
        r.records = records ? records[i] : nullptr;
        r.n_records = n_records ? n_records + i : nullptr;
        if (mdl_status s = r.begin())
            return s;
    }
    // step (or captured chunk) k of every driver, then the next: the streams
    // run side by side; where every driver can take the same chunk, one joint
    // graph launch advances them all
    for (bool more = true; more;) {
        more = false;
        if (n > 1 && joint_fits(plans)) {
            if (mdl_status s = joint_step(plans))
                return s;
            more = true;
            continue;
        }
        for (RunPlan &r : plans) {
            if (r.k >= nsteps)
                continue;
            more = true;
            MDL_HIP(hipSetDevice(r.d->device));
            if (mdl_status s = r.step())
                return s;
        }
    }
    mdl_status first = MDL_OK;
    for (RunPlan &r : plans) {
        mdl_status s = r.end();
        if (s != MDL_OK && first == MDL_OK)
            first = s;
    }
    return first;
}

mdl_status mdl_replicas_run(mdl_driver *const *drivers, int32_t n, int64_t nsteps,
                            int64_t sample_stride, int64_t rebuild_every, double *const *records,
                            int64_t *n_records) {
    if (!drivers || n < 1)
        return bad("no drivers");
    std::vector<Driver *> ds((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t j = 0; j < i; ++j)
            if (drivers[j] == drivers[i])
                return bad("a driver appears twice in mdl_replicas_run");
        ds[(size_t)i] = D(drivers[i]);
    }
    if (mdl_status s = replicas_check(ds))
        return s;
    std::vector<RunPlan> plans((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        RunPlan &r = plans[(size_t)i];
        r.d = D(drivers[i]);
        r.nsteps = nsteps;
        r.stride = sample_stride;
        r.rebuild = rebuild_every;
        r.records = records ? records[i] : nullptr;
        r.n_records = n_records ? n_records + i : nullptr;
        if (mdl_status s = r.begin())
            return s;
    }
    MDL_HIP(hipSetDevice(plans[0].d->device));
    mdl_status const s = replicas_run(plans, nsteps);
    mdl_status first = s;
    for (RunPlan &r : plans) {
        mdl_status const e = r.end();
        if (e != MDL_OK && first == MDL_OK)
            first = e;
    }
    return first;
}

mdl_status mdl_driver_set_graph_steps(mdl_driver *p, int32_t steps) {
    Driver *d = D(p);
    if (!d || steps < 0)
        return bad("mdl_driver_set_graph_steps: bad arguments");
    MDL_HIP(hipSetDevice(d->device));
    d->drop_graph();
    d->rep.drop_graph();
    joint_forget(d->uid);
    d->graph_steps = steps;
    return MDL_OK;
}

mdl_status mdl_driver_calculate_neighbors(mdl_driver *p) {
    Driver *d = D(p);
    if (mdl_status s = guard(d))
        return s;
    d->build_pending = true;   // the next step rebuilds the list after its drift
    return MDL_OK;
}

mdl_status mdl_driver_set_sample_observables(mdl_driver *p, int32_t flag) {
    Driver *d = D(p);
    if (!d)
        return bad("null driver");
    d->sample = flag ? 1 : 0;
    return MDL_OK;
}

mdl_status mdl_driver_get_sample_observables(mdl_driver *p, int32_t *flag) {
    Driver *d = D(p);
    if (!d || !flag)
        return bad("null argument");
    *flag = d->sample;
    return MDL_OK;
}

mdl_status mdl_driver_get_scalars(mdl_driver *p, double *out) {
    Driver *d = D(p);
    if (!d || !out)
        return bad("null argument");
    MDL_HIP(hipSetDevice(d->device));
    if (mdl_status s = d->sync_state())
        return s;
    memcpy(out, d->hs.rec, sizeof(double) * MDL_NSCALARS);
    out[MDL_S_NLIST_BUILDS] = (double)d->hs.nbuilds;
    out[MDL_S_WARN_VIRIAL] = (double)d->hs.warn_virial;
    out[MDL_S_WARN_RMIN] = (double)d->hs.warn_rmin;
    for (int k = 0; k < 3; ++k)
        out[MDL_S_BOX_INITIAL0 + k] = d->hs.L0[k];
    out[MDL_S_NLIST_ENTRIES] = (double)d->hs.nlist_entries;
    out[MDL_S_NMID_BUILDS] = (double)d->hs.nmid_builds;
    out[MDL_S_LANES_PER_ATOM] = (double)d->lpa;
    return MDL_OK;
}

mdl_status mdl_driver_get_atoms(mdl_driver *p, int32_t what, double *out) {
    Driver *d = D(p);
    if (!d || !out)
        return bad("null argument");
    if (what < MDL_A_POSITIONS || what > MDL_A_PAIRS_MID)
        return bad("unknown per-atom array");
    MDL_HIP(hipSetDevice(d->device));
    if (what == MDL_A_PAIRS_RC || what == MDL_A_PAIRS_MID) {
        if (!d->has_eam)
            return bad("MDL_A_PAIRS_RC / MDL_A_PAIRS_MID: the driver has no density pass");
        if (what == MDL_A_PAIRS_RC && d->precision == MDL_MIXED)
            k_count_within_rc<<<(d->N + 255) / 256, 256, 0, d->stream>>>(
                d->a, d->l, d->N, d->gs, (float)(d->rc2 * 1e16), d->d_stage);
        else
            k_int_to_double<<<(d->N + 255) / 256, 256, 0, d->stream>>>(
                what == MDL_A_PAIRS_RC ? d->l.ccnt : d->l.mcnt, d->N, d->d_stage);
        MDL_HIP(hipGetLastError());
        MDL_HIP(hipMemcpyAsync(out, d->d_stage, sizeof(double) * d->N, hipMemcpyDeviceToHost,
                               d->stream));
        MDL_HIP(hipStreamSynchronize(d->stream));
        return MDL_OK;
    }
    if (what == MDL_A_POSITIONS) {
        if (mdl_status s = d->sync_state())
            return s;
        if (d->hs.cycle == 0) {   // the ctor keeps the given positions (md_driver.cpp:31)
            memcpy(out, d->init_pos.data(), sizeof(double) * d->init_pos.size());
            return MDL_OK;
        }
    }
    int const n = (what <= MDL_A_FORCES) ? 3 * d->N : d->N;
    k_materialize<<<(d->N + 255) / 256, 256, 0, d->stream>>>(d->a, d->N, d->gs, what, d->d_stage);
    MDL_HIP(hipGetLastError());
    MDL_HIP(hipMemcpyAsync(out, d->d_stage, sizeof(double) * n, hipMemcpyDeviceToHost, d->stream));
    MDL_HIP(hipStreamSynchronize(d->stream));
    return MDL_OK;
}

mdl_status mdl_driver_cv_descriptors(mdl_driver *p, float rc, float mu0, float dmu, float width,
                                     float *Dp, float *Jp, float *Lp, float *La, double *pos_out,
                                     double *force_out, void *stream) {
    Driver *d = D(p);
    if (!d || !Dp || !Jp || !(rc > 0.f) || !(width > 0.f))
        return bad("mdl_driver_cv_descriptors: bad arguments");
    if (d->ntypes > 2)
        return bad("mdl_driver_cv_descriptors: the descriptors have two species channels");
    // lengths of the driver are in cm; the list holds every pair within hs.rc
    // (a float rc equal to the potential's cut-off rounds up by < 1e-7)
    if (!((double)rc * 1e-8 <= d->hs.rc * (1.0 + 1e-6)))
        return bad("mdl_driver_cv_descriptors: rc exceeds the neighbour list's cut-off");
    if (d->cycle_host == 0 && pos_out)
        return bad("mdl_driver_cv_descriptors: no step taken yet");
    MDL_HIP(hipSetDevice(d->device));
    if (d->failed)
        return bad("the driver stopped on an earlier error");
    // on the caller's stream: after the driver's queued work, and the driver's
    // later work after these reads; else on the driver's stream, synchronously
    hipStream_t const st = stream ? (hipStream_t)stream : d->stream;
    if (stream) {
        for (auto &e : d->ev_cv)
            if (!e)
                MDL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        MDL_HIP(hipEventRecord(d->ev_cv[0], d->stream));
        MDL_HIP(hipStreamWaitEvent(st, d->ev_cv[0], 0));
    }
    // the compacted pairs lie within the potential's cut-off rc2 (cm^2); a float
    // rc equal to it rounds up by < 1e-7
    int const which =
        (d->cycle_host == 0 || (double)rc * rc * 1e-16 > d->rc2 * (1.0 + 1e-6))
            ? 1
            : (d->precision == MDL_MIXED ? 2 : 0);
    int const waves = (d->N + 1) / 2;
    if (La)
        k_cv_desc_list<true><<<(waves * 64 + 255) / 256, 256, 0, st>>>(
            d->a, d->l, d->N, d->gs, 1e8, cv_basis(rc, mu0, dmu, width), which, Dp, Jp, Lp,
            La);
    else
        k_cv_desc_list<false><<<(waves * 64 + 255) / 256, 256, 0, st>>>(
            d->a, d->l, d->N, d->gs, 1e8, cv_basis(rc, mu0, dmu, width), which, Dp, Jp, Lp,
            nullptr);
    MDL_HIP(hipGetLastError());
    unsigned const g = (d->N + 255) / 256;
    if (pos_out) {
        k_materialize<<<g, 256, 0, st>>>(d->a, d->N, d->gs, MDL_A_POSITIONS, pos_out);
        MDL_HIP(hipGetLastError());
    }
    if (force_out) {
        k_materialize<<<g, 256, 0, st>>>(d->a, d->N, d->gs, MDL_A_FORCES, force_out);
        MDL_HIP(hipGetLastError());
    }
    if (stream) {
        MDL_HIP(hipEventRecord(d->ev_cv[1], st));
        MDL_HIP(hipStreamWaitEvent(d->stream, d->ev_cv[1], 0));
    } else {
        MDL_HIP(hipStreamSynchronize(d->stream));
    }
    return MDL_OK;
}

mdl_status mdl_driver_pair_correlation(mdl_driver *p, const double *rg, int32_t nbin,
                                       int32_t mode, double *gr, double *gr_partial) {
    Driver *d = D(p);
    GrCall c;
    if (mdl_status s = gr_prepare(d, rg, nbin, mode, gr, gr_partial, c))
        return s;
    unsigned long long *cnt = nullptr;
    MDL_HIP(hipMalloc(&cnt, c.nh * sizeof(unsigned long long)));
    std::vector<unsigned long long> h(c.nh);
    hipError_t e = gr_count(d, c, cnt);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), cnt, c.nh * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(d->stream);
    (void)hipFree(cnt);
    if (e != hipSuccess)
        return hip_fail(e, "mdl_driver_pair_correlation");
    for (int row = 0; row < c.nrow; ++row)
        if (double *out = gr_row_out(c, row, gr, gr_partial))
            gr_normalize(c, row, 0, nbin, h.data() + (size_t)row * nbin, rg, out + 0);
    return MDL_OK;
}

mdl_status mdl_driver_pair_correlation_noalloc(mdl_driver *p, const double *rg, int32_t nbin,
                                               int32_t mode, double *gr, double *gr_partial) {
    Driver *d = D(p);
    GrCall c;
    if (mdl_status s = gr_prepare(d, rg, nbin, mode, gr, gr_partial, c,
                                  (size_t)MDL_GR_SCRATCH_COUNTS))
        return s;
    std::lock_guard<std::mutex> lock(g_gr_scratch_mutex);
    unsigned long long *cnt = nullptr;   // device address of `scratch` on this device
    if (mdl_status s = gr_scratch(d->device, &cnt))
        return s;
    hipError_t e = gr_count(d, c, cnt);
    // counts to the host through a fixed buffer, one chunk of a row at a time
    constexpr int CHUNK = 512;
    unsigned long long buf[CHUNK];
    for (int row = 0; row < c.nrow && e == hipSuccess; ++row) {
        double *out = gr_row_out(c, row, gr, gr_partial);
        if (!out)
            continue;
        for (int k0 = 0; k0 < nbin && e == hipSuccess; k0 += CHUNK) {
            int const n = std::min(CHUNK, nbin - k0);
            e = hipMemcpyAsync(buf, cnt + (size_t)row * nbin + k0,
                               (size_t)n * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               d->stream);
            if (e == hipSuccess)
                e = hipStreamSynchronize(d->stream);
            if (e == hipSuccess)
                gr_normalize(c, row, k0, n, buf, rg, out);
        }
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess)
        return hip_fail(e, "mdl_driver_pair_correlation_noalloc");
    return MDL_OK;
}

mdl_status mdl_driver_set_cv_capture(mdl_driver *p, float rc, float mu0, float dmu, float width,
                                     float *Dp, float *Jp, float *Lp, double *pos_out,
                                     double *force_out, int32_t slots) {
    Driver *d = D(p);
    if (!d || slots < 0)
        return bad("mdl_driver_set_cv_capture: bad arguments");
    if (slots == 0) {
        d->cvc = Driver::CvCapture{};
        return MDL_OK;
    }
    if (!Dp || !Jp || !Lp || !pos_out || !force_out || !(rc > 0.f) || !(width > 0.f))
        return bad("mdl_driver_set_cv_capture: bad arguments");
    if (d->ntypes > 2)
        return bad("mdl_driver_set_cv_capture: the descriptors have two species channels");
    if (!((double)rc * 1e-8 <= d->hs.rc * (1.0 + 1e-6)))
        return bad("mdl_driver_set_cv_capture: rc exceeds the neighbour list's cut-off");
    Driver::CvCapture c;
    c.D = Dp;
    c.J = Jp;
    c.L = Lp;
    c.pos = pos_out;
    c.force = force_out;
    c.slots = slots;
    c.next = 0;
    c.list = (double)rc * rc * 1e-16 > d->rc2 * (1.0 + 1e-6)
                     ? 1
                     : (d->precision == MDL_MIXED ? 2 : 0);
    c.basis = cv_basis(rc, mu0, dmu, width);
    d->cvc = c;
    return MDL_OK;
}

mdl_status mdl_driver_set_cv_capture_pairs(mdl_driver *p, int32_t *nbr_out, int32_t *count_out,
                                           int32_t width) {
    Driver *d = D(p);
    if (!d || (nbr_out && (!count_out || width < 1)))
        return bad("mdl_driver_set_cv_capture_pairs: bad arguments");
    if (nbr_out && d->cvc.slots == 0)
        return bad("mdl_driver_set_cv_capture_pairs: no capture is set");
    d->cvc.nbr = nbr_out;
    d->cvc.ncnt = nbr_out ? count_out : nullptr;
    d->cvc.width = nbr_out ? width : 0;
    return MDL_OK;
}

mdl_status mdl_driver_stream(mdl_driver *p, void **out) {
    Driver *d = D(p);
    if (!d || !out)
        return bad("null argument");
    *out = (void *)d->stream;
    return MDL_OK;
}

mdl_status mdl_driver_copy_atoms(mdl_driver *p, int32_t what, double *out) {
    Driver *d = D(p);
    if (!d || !out)
        return bad("null argument");
    if (what < MDL_A_POSITIONS || what > MDL_A_EKIN)
        return bad("unknown per-atom array");
    MDL_HIP(hipSetDevice(d->device));
    int const n = (what <= MDL_A_FORCES) ? 3 * d->N : d->N;
    if (what == MDL_A_POSITIONS && d->cycle_host == 0) {
        // the ctor keeps the given positions (md_driver.cpp:31)
        MDL_HIP(hipMemcpyAsync(out, d->init_pos.data(), sizeof(double) * n,
                               hipMemcpyHostToDevice, d->stream));
    } else {
        k_materialize<<<(d->N + 255) / 256, 256, 0, d->stream>>>(d->a, d->N, d->gs, what, out);
        MDL_HIP(hipGetLastError());
    }
    MDL_HIP(hipStreamSynchronize(d->stream));
    return MDL_OK;
}

mdl_status mdl_driver_get_chain(mdl_driver *p, int32_t which, double *p_eta, double *eta) {
    Driver *d = D(p);
    if (!d || !p_eta || !eta)
        return bad("null argument");
    MDL_HIP(hipSetDevice(d->device));
    if (mdl_status s = d->sync_state())
        return s;
    const Chain &c = which ? d->hs.bth : d->hs.th;
    memcpy(p_eta, c.p_eta, sizeof(double) * c.M);
    memcpy(eta, c.eta, sizeof(double) * c.M);
    return MDL_OK;
}

mdl_status mdl_driver_kernel_times(mdl_driver *p, double *out) {
    Driver *d = D(p);
    if (!d || !out)
        return bad("null argument");
    out[0] = d->t_dens;
    out[1] = d->t_force;
    out[2] = d->t_step;
    return MDL_OK;
}

mdl_status mdl_expx_eval(int32_t form, int64_t n, const double *x, double *expx, double *op) {
    if (form < 0 || form > 2 || n < 0 || (n > 0 && (!x || !expx || !op)))
        return bad("mdl_expx_eval: bad argument");
    if (n == 0)
        return MDL_OK;
    double *buf = nullptr;
    MDL_HIP(hipMalloc(&buf, sizeof(double) * 3 * (size_t)n));
    double *dx = buf, *de = buf + n, *dop = buf + 2 * n;
    hipError_t e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_expx_eval<<<(unsigned)((n + BLOCK - 1) / BLOCK), BLOCK>>>(form, n, dx, de, dop);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(expx, de, sizeof(double) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(op, dop, sizeof(double) * n, hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (e != hipSuccess)
        return hip_fail(e, "mdl_expx_eval");
    return MDL_OK;
}

}  // extern "C"
