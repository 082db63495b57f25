// Compiled UCC programs: energy, adjoint gradient and pool gradients with the
// ansatz and the Hamiltonian resident in HBM.
//
// The reference rebuilds the ansatz circuit (JW of every pool operator,
// trotterize, gate objects) on every energy_feval (algorithm.py:458-474,
// ansatz.py:25-82) and evaluates the gradient with deep copies and Python
// list conversions (uccvqeabc.py:122-251).  A program is compiled once; the
// amplitudes are the only per-call input.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "qfe_device.h"
#include "qfe_internal.h"

namespace qfe {

// A grouped Pauli table resident in HBM.
struct DevTable {
    PauliGroup* g = nullptr;
    uint64_t* z = nullptr;
    double2* c = nullptr;
    uint32_t ngroups = 0;
};

static int table_upload(const std::vector<PauliGroup>& groups, const std::vector<uint64_t>& zs,
                        const std::vector<double2>& cs, DevTable& t) {
    t.ngroups = (uint32_t)groups.size();
    if (groups.empty()) return QFE_OK;
    QFE_HIP(hipMalloc(&t.g, groups.size() * sizeof(PauliGroup)));
    QFE_HIP(hipMalloc(&t.z, zs.size() * sizeof(uint64_t)));
    QFE_HIP(hipMalloc(&t.c, cs.size() * sizeof(double2)));
    QFE_HIP(hipMemcpy(t.g, groups.data(), groups.size() * sizeof(PauliGroup), hipMemcpyHostToDevice));
    QFE_HIP(hipMemcpy(t.z, zs.data(), zs.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    QFE_HIP(hipMemcpy(t.c, cs.data(), cs.size() * sizeof(double2), hipMemcpyHostToDevice));
    return QFE_OK;
}

static void table_free(DevTable& t) {
    if (t.g) (void)hipFree(t.g);
    if (t.z) (void)hipFree(t.z);
    if (t.c) (void)hipFree(t.c);
    t = DevTable{};
}

// Concatenate the groupings of several operators; ranges[k]..ranges[k+1] are
// the groups of operator k.
static void group_many(const qfe_pauli* terms, const uint64_t* offsets, uint64_t nops, std::vector<PauliGroup>& groups,
                       std::vector<uint64_t>& zs, std::vector<double2>& cs, std::vector<uint32_t>& ranges) {
    groups.clear();
    zs.clear();
    cs.clear();
    ranges.assign(1, 0);
    std::vector<PauliGroup> g1;
    std::vector<uint64_t> z1;
    std::vector<double2> c1;
    for (uint64_t k = 0; k < nops; ++k) {
        group_terms(terms + offsets[k], offsets[k + 1] - offsets[k], g1, z1, c1);
        const uint32_t shift = (uint32_t)zs.size();
        for (PauliGroup g : g1) {
            g.first += shift;
            groups.push_back(g);
        }
        zs.insert(zs.end(), z1.begin(), z1.end());
        cs.insert(cs.end(), c1.begin(), c1.end());
        ranges.push_back((uint32_t)groups.size());
    }
}

// ---- captured energy sequence --------------------------------------------------------------
// The ansatz is planned once with device-side angles: every exponential's
// (cos, sin) lives in a descriptor image in HBM and a kernel rewrites them
// from the amplitudes before the passes run.  Parameter upload, that kernel,
// the reference state, the exponential passes, the expectation value and the
// result download are captured once as a hipGraph; each energy call then
// writes the amplitudes into pinned memory and replays the graph.
struct DevAngle {
    uint64_t off;  // byte offset of the (cos, sin) pair in the descriptor image
    uint32_t param;
    uint32_t pad;
    double scale;  // generator scale
    double cim;    // Im(coefficient) of the Pauli term
};

__global__ __launch_bounds__(256) void k_program_angles(const double* __restrict__ params,
                                                        const DevAngle* __restrict__ a, uint32_t n,
                                                        char* __restrict__ image) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevAngle d = a[i];
    double sn, c;
    sincos((params[d.param] * d.scale) * d.cim, &sn, &c);
    double* dst = reinterpret_cast<double*>(image + d.off);  // 8-byte aligned in every descriptor kind
    dst[0] = c;
    dst[1] = sn;
}

struct ProgramGraph {
    OpPlan* plan = nullptr;
    char* image = nullptr;
    DevAngle* angles = nullptr;
    uint32_t nangles = 0;
    double* dparams = nullptr;
    double* hparams = nullptr;  // pinned; read by the graph's first node
    double2* partial = nullptr;
    double2* gout = nullptr;
    double2* hout = nullptr;  // pinned; written by the graph's last node
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void graph_free(ProgramGraph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->image) (void)hipFree(g->image);
    if (g->angles) (void)hipFree(g->angles);
    if (g->dparams) (void)hipFree(g->dparams);
    if (g->partial) (void)hipFree(g->partial);
    if (g->gout) (void)hipFree(g->gout);
    if (g->hparams) (void)hipHostFree(g->hparams);
    if (g->hout) (void)hipHostFree(g->hout);
    plan_free(g->plan);
    delete g;
}

}  // namespace qfe

struct qfe_program_s {
    int n = 0;
    uint64_t ref = 0;
    std::vector<qfe_pauli> gen_terms;
    std::vector<uint64_t> gen_off;
    std::vector<int32_t> gparam;
    std::vector<double> gscale;
    uint64_t nparam = 0;
    std::vector<qfe_pauli> ham_terms;
    qfe::DevTable ham;      // Hamiltonian grouped by X mask
    qfe::DevTable gens;     // every generator grouped by X mask
    std::vector<uint32_t> gen_ranges;
    double2* partial = nullptr;  // reduction partials (reused launch to launch)
    size_t partial_cap = 0;
    double2* gout = nullptr;     // per-group results (generators, then Hamiltonian)
    size_t gout_cap = 0;
    qfe_state psi = nullptr;
    qfe_state sig = nullptr;
    qfe::ProgramGraph* graph = nullptr;  // captured energy sequence (first graph energy call)
};

namespace qfe {

static int ensure_partial(qfe_program p, size_t entries) {
    if (p->partial_cap >= entries) return QFE_OK;
    if (p->partial) {
        QFE_HIP(hipStreamSynchronize(ctx().stream));
        QFE_HIP(hipFree(p->partial));
        p->partial = nullptr;
    }
    QFE_HIP(hipMalloc(&p->partial, entries * sizeof(double2)));
    p->partial_cap = entries;
    return QFE_OK;
}

static int ensure_gout(qfe_program p, size_t entries) {
    if (p->gout_cap >= entries) return QFE_OK;
    if (p->gout) {
        QFE_HIP(hipStreamSynchronize(ctx().stream));
        QFE_HIP(hipFree(p->gout));
        p->gout = nullptr;
    }
    QFE_HIP(hipMalloc(&p->gout, entries * sizeof(double2)));
    p->gout_cap = entries;
    return QFE_OK;
}

static int ensure_states(qfe_program p, bool need_sigma) {
    int rc = QFE_OK;
    if (!p->psi) rc = qfe_state_create(p->n, &p->psi);
    if (rc) return rc;
    if (need_sigma && !p->sig) rc = qfe_state_create(p->n, &p->sig);
    return rc;
}

// Lower generator k (forward: sign = +1; inverse: sign = -1, terms reversed).
static int lower_generator(qfe_program p, uint64_t k, const double* params, double sign, std::vector<HostOp>& ops) {
    const double theta = params[p->gparam[k]] * p->gscale[k] * sign;
    const uint64_t b = p->gen_off[k], e = p->gen_off[k + 1];
    for (uint64_t i = 0; i < e - b; ++i) {
        const qfe_pauli& t = p->gen_terms[sign > 0 ? b + i : e - 1 - i];
        // exp(theta * coeff * P) with coeff purely imaginary = exp(i * theta * Im(coeff) * P)
        int rc = lower_pexp(t, theta * t.coeff.im, p->n, ops);
        if (rc) return rc;
    }
    return QFE_OK;
}

static int prepare(qfe_program p, const double* params, qfe_state out) {
    int rc = launch_basis(out->d, out->N, p->ref, ctx().stream);
    if (rc) return rc;
    std::vector<HostOp> ops;
    ops.reserve(p->gen_terms.size());
    for (uint64_t k = 0; k + 1 < p->gen_off.size(); ++k) {
        rc = lower_generator(p, k, params, +1.0, ops);
        if (rc) return rc;
    }
    return run_ops(out, ops);
}

static int fetch(const double2* dev, size_t count, std::vector<double2>& host) {
    host.resize(count);
    if (!count) return QFE_OK;
    void* pin = nullptr;
    int rc = pinned_reserve(count * sizeof(double2), &pin);
    if (rc) return rc;
    QFE_HIP(hipMemcpyAsync(pin, dev, count * sizeof(double2), hipMemcpyDeviceToHost, ctx().stream));
    QFE_HIP(hipStreamSynchronize(ctx().stream));
    std::memcpy(host.data(), pin, count * sizeof(double2));
    return QFE_OK;
}

static int energy_of(qfe_program p, qfe_state s, double* energy) {
    const uint32_t nch = expval_chunks(p->n, p->ham.ngroups);
    int rc = ensure_partial(p, (size_t)p->ham.ngroups * nch);
    if (rc) return rc;
    rc = ensure_gout(p, p->ham.ngroups);
    if (rc) return rc;
    rc = launch_expval_async(s->d, s->d, p->n, p->ham.g, p->ham.ngroups, p->ham.z, p->ham.c, p->partial, nch, p->gout);
    if (rc) return rc;
    std::vector<double2> per;
    rc = fetch(p->gout, p->ham.ngroups, per);
    if (rc) return rc;
    double e = 0.0;
    for (const double2& v : per) e += v.x;
    *energy = e;
    return QFE_OK;
}

// Plan the ansatz with device-side angles and capture the energy sequence.
static int graph_build(qfe_program p) {
    int rc = ensure_states(p, false);
    if (rc) return rc;
    // every term with a nonzero coefficient becomes an op, whatever the amplitudes
    std::vector<HostOp> ops;
    std::vector<DevAngle> table;
    for (uint64_t k = 0; k + 1 < p->gen_off.size(); ++k) {
        if (p->gscale[k] == 0.0) continue;
        for (uint64_t i = p->gen_off[k]; i < p->gen_off[k + 1]; ++i) {
            const qfe_pauli& t = p->gen_terms[i];
            if (t.coeff.im == 0.0) continue;
            rc = lower_pexp(t, 1.0, p->n, ops);
            if (rc) return rc;
            ops.back().src = (int32_t)table.size();
            table.push_back(DevAngle{0, (uint32_t)p->gparam[k], 0, p->gscale[k], t.coeff.im});
        }
    }
    ProgramGraph* g = new ProgramGraph();
    std::vector<char> image;
    std::vector<AngleRef> refs;
    rc = plan_build(ops, p->n, true, &g->plan, image, refs);
    if (rc) {
        graph_free(g);
        return rc;
    }
    std::vector<DevAngle> angles;
    angles.reserve(refs.size());
    for (const AngleRef& r : refs) {
        DevAngle a = table[r.src];
        a.off = r.off;
        angles.push_back(a);
    }
    g->nangles = (uint32_t)angles.size();
    const uint32_t nhg = p->ham.ngroups;
    const uint32_t nch = expval_chunks(p->n, nhg);
    auto alloc = [&]() -> int {
        if (!image.empty()) QFE_HIP(hipMalloc(&g->image, image.size()));
        if (!angles.empty()) QFE_HIP(hipMalloc(&g->angles, angles.size() * sizeof(DevAngle)));
        QFE_HIP(hipMalloc(&g->dparams, (p->nparam + 1) * sizeof(double)));
        QFE_HIP(hipMalloc(&g->partial, ((size_t)nhg * nch + 1) * sizeof(double2)));
        QFE_HIP(hipMalloc(&g->gout, ((size_t)nhg + 1) * sizeof(double2)));
        QFE_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->hparams), (p->nparam + 1) * sizeof(double), hipHostMallocDefault));
        QFE_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->hout), ((size_t)nhg + 1) * sizeof(double2), hipHostMallocDefault));
        if (!image.empty()) QFE_HIP(hipMemcpy(g->image, image.data(), image.size(), hipMemcpyHostToDevice));
        if (!angles.empty())
            QFE_HIP(hipMemcpy(g->angles, angles.data(), angles.size() * sizeof(DevAngle), hipMemcpyHostToDevice));
        return QFE_OK;
    };
    rc = alloc();
    if (rc) {
        graph_free(g);
        return rc;
    }
    hipStream_t st = ctx().stream;
    QFE_HIP(hipStreamSynchronize(st));
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        graph_free(g);
        return hip_fail(e, "qfe_program_energy_graph: begin capture");
    }
    // launches below only record; the capture is always ended before returning
    auto record = [&]() -> int {
        if (p->nparam)
            QFE_HIP(hipMemcpyAsync(g->dparams, g->hparams, p->nparam * sizeof(double), hipMemcpyHostToDevice, st));
        if (g->nangles) {
            hipLaunchKernelGGL(k_program_angles, dim3((g->nangles + 255) / 256), dim3(256), 0, st, g->dparams, g->angles,
                               g->nangles, g->image);
            QFE_HIP(hipGetLastError());
        }
        int r = launch_basis(p->psi->d, p->psi->N, p->ref, st);
        if (!r) r = plan_launch(g->plan, p->psi, g->image);
        if (!r) r = launch_expval_async(p->psi->d, p->psi->d, p->n, p->ham.g, nhg, p->ham.z, p->ham.c, g->partial, nch, g->gout);
        if (r) return r;
        if (nhg) QFE_HIP(hipMemcpyAsync(g->hout, g->gout, nhg * sizeof(double2), hipMemcpyDeviceToHost, st));
        return QFE_OK;
    };
    rc = record();
    e = hipStreamEndCapture(st, &g->graph);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "qfe_program_energy_graph: end capture");
    if (!rc) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "qfe_program_energy_graph: instantiate");
    }
    if (rc) {
        graph_free(g);
        return rc;
    }
    p->graph = g;
    return QFE_OK;
}

static int graph_energy(qfe_program p, const double* params, double* energy) {
    ProgramGraph* g = p->graph;
    hipStream_t st = ctx().stream;
    // the previous replay has finished (every call synchronizes), so the pinned buffers are free
    if (p->nparam) std::memcpy(g->hparams, params, p->nparam * sizeof(double));
    QFE_HIP(hipGraphLaunch(g->exec, st));
    QFE_HIP(hipStreamSynchronize(st));
    double e = 0.0;
    for (uint32_t i = 0; i < p->ham.ngroups; ++i) e += g->hout[i].x;
    *energy = e;
    return QFE_OK;
}

int program_nqubit(qfe_program p) { return p->n; }
uint64_t program_nparam(qfe_program p) { return p->nparam; }
ProgramView program_view(qfe_program p) {
    return ProgramView{p->n, p->ref, p->nparam, &p->gen_terms, &p->gen_off, &p->gparam, &p->gscale, &p->ham_terms};
}

}  // namespace qfe

using namespace qfe;

extern "C" {

int qfe_program_create(int nqubit, uint64_t ref_index, const qfe_pauli* gen_terms, const uint64_t* gen_offsets,
                       uint64_t ngen, const int32_t* gen_param, const double* gen_scale, uint64_t nparam,
                       const qfe_pauli* ham_terms, uint64_t nham, qfe_program* out) {
    if (!out) return fail(QFE_ERR_ARG, "qfe_program_create: null output");
    *out = nullptr;
    if (nqubit < 1 || nqubit > 36) return fail(QFE_ERR_ARG, "qfe_program_create: nqubit must be in [1, 36]");
    if (ref_index >> nqubit) return fail(QFE_ERR_ARG, "qfe_program_create: reference index out of range");
    if (ngen && (!gen_terms || !gen_offsets || !gen_param || !gen_scale))
        return fail(QFE_ERR_ARG, "qfe_program_create: null generator arrays");
    if (nham && !ham_terms) return fail(QFE_ERR_ARG, "qfe_program_create: null Hamiltonian");
    const uint64_t lim = (1ull << nqubit) - 1ull;
    if (ngen && gen_offsets[0] != 0) return fail(QFE_ERR_ARG, "qfe_program_create: gen_offsets[0] must be 0");
    for (uint64_t k = 0; k < ngen; ++k) {
        if (gen_offsets[k + 1] < gen_offsets[k]) return fail(QFE_ERR_ARG, "qfe_program_create: offsets must be non-decreasing");
        if (gen_param[k] < 0 || (uint64_t)gen_param[k] >= nparam)
            return fail(QFE_ERR_ARG, "qfe_program_create: generator parameter index out of range");
    }
    const uint64_t nterms = ngen ? gen_offsets[ngen] : 0;
    for (uint64_t i = 0; i < nterms; ++i) {
        const qfe_pauli& t = gen_terms[i];
        if ((t.x | t.z) & ~lim) return fail(QFE_ERR_ARG, "qfe_program_create: generator term acts outside the register");
        if (std::fabs(t.coeff.re) > 1.0e-14)
            return fail(QFE_ERR_ARG, "qfe_program_create: generator coefficients must be purely imaginary");
    }
    for (uint64_t i = 0; i < nham; ++i)
        if ((ham_terms[i].x | ham_terms[i].z) & ~lim) return fail(QFE_ERR_ARG, "qfe_program_create: Hamiltonian term acts outside the register");
    int rc = ensure_init();
    if (rc) return rc;
    qfe_program p = new qfe_program_s();
    p->n = nqubit;
    p->ref = ref_index;
    p->gen_terms.assign(gen_terms, gen_terms + nterms);
    p->gen_off.assign(gen_offsets, gen_offsets + (ngen ? ngen + 1 : 1));
    if (!ngen) p->gen_off[0] = 0;
    p->gparam.assign(gen_param, gen_param + ngen);
    p->gscale.assign(gen_scale, gen_scale + ngen);
    p->nparam = nparam;
    p->ham_terms.assign(ham_terms, ham_terms + nham);
    std::vector<PauliGroup> g;
    std::vector<uint64_t> z;
    std::vector<double2> c;
    group_terms(ham_terms, nham, g, z, c);
    rc = table_upload(g, z, c, p->ham);
    if (!rc && ngen) {
        group_many(gen_terms, gen_offsets, ngen, g, z, c, p->gen_ranges);
        rc = table_upload(g, z, c, p->gens);
    }
    if (rc) {
        qfe_program_destroy(p);
        return rc;
    }
    *out = p;
    return QFE_OK;
}

int qfe_program_destroy(qfe_program p) {
    if (!p) return QFE_OK;
    (void)hipStreamSynchronize(ctx().stream);
    table_free(p->ham);
    table_free(p->gens);
    if (p->partial) (void)hipFree(p->partial);
    if (p->gout) (void)hipFree(p->gout);
    graph_free(p->graph);
    qfe_state_destroy(p->psi);
    qfe_state_destroy(p->sig);
    delete p;
    return QFE_OK;
}

int qfe_program_prepare(qfe_program p, const double* params, qfe_state out) {
    if (!p || !out || !out->d || (p->nparam && !params)) return fail(QFE_ERR_ARG, "qfe_program_prepare: invalid argument");
    if (out->n != p->n) return fail(QFE_ERR_ARG, "qfe_program_prepare: state has the wrong number of qubits");
    return prepare(p, params, out);
}

int qfe_program_energy(qfe_program p, const double* params, double* energy) {
    if (!p || !energy || (p->nparam && !params)) return fail(QFE_ERR_ARG, "qfe_program_energy: invalid argument");
    int rc = ensure_states(p, false);
    if (rc) return rc;
    rc = prepare(p, params, p->psi);
    if (rc) return rc;
    return energy_of(p, p->psi, energy);
}

int qfe_program_energy_graph(qfe_program p, const double* params, double* energy) {
    if (!p || !energy || (p->nparam && !params)) return fail(QFE_ERR_ARG, "qfe_program_energy_graph: invalid argument");
    if (!p->graph) {
        int rc = graph_build(p);
        if (rc) return rc;
    }
    return graph_energy(p, params, energy);
}

int qfe_program_gradient(qfe_program p, const double* params, double* grad, double* energy) {
    if (!p || !grad || (p->nparam && !params)) return fail(QFE_ERR_ARG, "qfe_program_gradient: invalid argument");
    int rc = ensure_states(p, true);
    if (rc) return rc;
    rc = prepare(p, params, p->psi);
    if (rc) return rc;
    const uint64_t ngen = p->gen_off.size() - 1;
    const uint32_t ngg = p->gens.ngroups;
    const uint32_t nhg = p->ham.ngroups;
    // sigma = H psi
    rc = launch_apply_sum_async(p->psi->d, p->sig->d, p->n, p->ham.g, nhg, p->ham.z, p->ham.c);
    if (rc) return rc;
    uint32_t maxg = nhg;
    for (uint64_t k = 0; k < ngen; ++k) maxg = std::max(maxg, p->gen_ranges[k + 1] - p->gen_ranges[k]);
    const uint32_t nch = expval_chunks(p->n, maxg);
    rc = ensure_partial(p, (size_t)maxg * nch);
    if (rc) return rc;
    rc = ensure_gout(p, (size_t)ngg + nhg);
    if (rc) return rc;
    if (energy) {
        rc = launch_expval_async(p->psi->d, p->psi->d, p->n, p->ham.g, nhg, p->ham.z, p->ham.c, p->partial, nch,
                                 p->gout + ngg);
        if (rc) return rc;
    }
    // reverse sweep: <sigma_k| A_k |psi_k>, then un-apply generator k on both
    for (uint64_t kk = ngen; kk-- > 0;) {
        const uint32_t g0 = p->gen_ranges[kk], g1 = p->gen_ranges[kk + 1];
        rc = launch_expval_async(p->sig->d, p->psi->d, p->n, p->gens.g + g0, g1 - g0, p->gens.z, p->gens.c, p->partial,
                                 nch, p->gout + g0);
        if (rc) return rc;
        if (kk == 0) break;
        std::vector<HostOp> ops;
        rc = lower_generator(p, kk, params, -1.0, ops);
        if (rc) return rc;
        std::vector<HostOp> ops2 = ops;
        rc = run_ops(p->psi, ops);
        if (rc) return rc;
        rc = run_ops(p->sig, ops2);
        if (rc) return rc;
    }
    std::vector<double2> per;
    rc = fetch(p->gout, (size_t)ngg + nhg, per);
    if (rc) return rc;
    for (uint64_t i = 0; i < p->nparam; ++i) grad[i] = 0.0;
    for (uint64_t k = 0; k < ngen; ++k) {
        double acc = 0.0;
        for (uint32_t g = p->gen_ranges[k]; g < p->gen_ranges[k + 1]; ++g) acc += per[g].x;
        grad[p->gparam[k]] += 2.0 * p->gscale[k] * acc;
    }
    if (energy) {
        double e = 0.0;
        for (uint32_t g = 0; g < nhg; ++g) e += per[ngg + g].x;
        *energy = e;
    }
    return QFE_OK;
}

int qfe_program_residuals(qfe_program p, const double* params, const uint64_t* idx, uint64_t count, qfe_c64* out) {
    if (!p || (count && (!idx || !out)) || (p->nparam && !params))
        return fail(QFE_ERR_ARG, "qfe_program_residuals: invalid argument");
    for (uint64_t i = 0; i < count; ++i)
        if (idx[i] >> p->n) return fail(QFE_ERR_ARG, "qfe_program_residuals: basis index out of range");
    int rc = ensure_states(p, true);
    if (rc) return rc;
    rc = prepare(p, params, p->psi);
    if (rc) return rc;
    // sigma = H U|ref>, then U^dag: generators last to first, terms reversed
    rc = launch_apply_sum_async(p->psi->d, p->sig->d, p->n, p->ham.g, p->ham.ngroups, p->ham.z, p->ham.c);
    if (rc) return rc;
    std::vector<HostOp> ops;
    for (uint64_t kk = p->gen_off.size() - 1; kk-- > 0;) {
        rc = lower_generator(p, kk, params, -1.0, ops);
        if (rc) return rc;
    }
    rc = run_ops(p->sig, ops);
    if (rc) return rc;
    for (uint64_t i = 0; i < count; ++i) {
        rc = qfe_state_get_amp(p->sig, idx[i], &out[i]);
        if (rc) return rc;
    }
    return QFE_OK;
}

int qfe_program_pool_gradients(qfe_program p, const double* params, const qfe_pauli* pool_terms,
                               const uint64_t* op_offsets, uint64_t nops, uint64_t batch, double* grads) {
    if (!p || (nops && (!pool_terms || !op_offsets || !grads)) || (p->nparam && !params))
        return fail(QFE_ERR_ARG, "qfe_program_pool_gradients: invalid argument");
    if (nops == 0) return QFE_OK;
    if (batch == 0) batch = 64;
    const uint64_t lim = (1ull << p->n) - 1ull;
    for (uint64_t i = 0; i < op_offsets[nops]; ++i)
        if ((pool_terms[i].x | pool_terms[i].z) & ~lim) return fail(QFE_ERR_ARG, "qfe_program_pool_gradients: term outside the register");
    int rc = ensure_states(p, true);
    if (rc) return rc;
    rc = prepare(p, params, p->psi);
    if (rc) return rc;
    rc = launch_apply_sum_async(p->psi->d, p->sig->d, p->n, p->ham.g, p->ham.ngroups, p->ham.z, p->ham.c);
    if (rc) return rc;
    std::vector<PauliGroup> g;
    std::vector<uint64_t> z;
    std::vector<double2> c;
    std::vector<uint32_t> ranges;
    group_many(pool_terms, op_offsets, nops, g, z, c, ranges);
    DevTable t;
    rc = table_upload(g, z, c, t);
    if (rc) {
        table_free(t);
        return rc;
    }
    rc = ensure_gout(p, g.size());
    // batches of `batch` operators per launch
    for (uint64_t b0 = 0; !rc && b0 < nops; b0 += batch) {
        const uint64_t b1 = std::min(nops, b0 + batch);
        const uint32_t g0 = ranges[b0], g1 = ranges[b1];
        if (g1 == g0) continue;
        const uint32_t nch = expval_chunks(p->n, g1 - g0);
        rc = ensure_partial(p, (size_t)(g1 - g0) * nch);
        if (!rc)
            rc = launch_expval_async(p->sig->d, p->psi->d, p->n, t.g + g0, g1 - g0, t.z, t.c, p->partial, nch,
                                     p->gout + g0);
    }
    std::vector<double2> per;
    if (!rc) rc = fetch(p->gout, g.size(), per);
    if (!rc) {
        for (uint64_t k = 0; k < nops; ++k) {
            double acc = 0.0;
            for (uint32_t gi = ranges[k]; gi < ranges[k + 1]; ++gi) acc += per[gi].x;
            grads[k] = 2.0 * acc;
        }
    }
    (void)hipStreamSynchronize(ctx().stream);
    table_free(t);
    return rc;
}

}  // extern "C"
