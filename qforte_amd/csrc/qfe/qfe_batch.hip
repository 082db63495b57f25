// Serving engine: continuous batching of UCC energy requests.
//
// Requests (a compiled program plus its amplitudes) queue on the host.  A
// request is admitted to an active slot when a slot is free and its state
// (16 * 2^n bytes) fits the HBM budget; every engine step advances all
// active requests and retires those that finished, so requests arriving while
// others run are admitted as soon as room frees up.
#include <hip/hip_runtime.h>

#include <deque>
#include <vector>

#include "qfe_common.cuh"
#include "qfe_internal.h"

namespace {

struct Request {
    uint64_t id;
    qfe_program prog;
    std::vector<double> params;
};

}  // namespace

struct qfe_batch_s {
    uint64_t budget = 0;
    int max_active = 0;
    uint64_t next_id = 1;
    uint64_t used = 0;
    std::deque<Request> queue;
    std::vector<Request> active;
};

using namespace qfe;

extern "C" {

int qfe_batch_create(uint64_t hbm_budget_bytes, int max_active, qfe_batch* out) {
    if (!out || max_active < 1 || hbm_budget_bytes == 0) return fail(QFE_ERR_ARG, "qfe_batch_create: invalid argument");
    int rc = ensure_init();
    if (rc) return rc;
    qfe_batch b = new qfe_batch_s();
    b->budget = hbm_budget_bytes;
    b->max_active = max_active;
    *out = b;
    return QFE_OK;
}

int qfe_batch_destroy(qfe_batch b) {
    delete b;
    return QFE_OK;
}

int qfe_batch_submit(qfe_batch b, qfe_program p, const double* params, uint64_t* request_id) {
    if (!b || !p || !request_id) return fail(QFE_ERR_ARG, "qfe_batch_submit: invalid argument");
    const uint64_t np = program_nparam(p);
    if (np && !params) return fail(QFE_ERR_ARG, "qfe_batch_submit: null parameters");
    const uint64_t foot = 16ull << program_nqubit(p);
    if (foot > b->budget) return fail(QFE_ERR_ARG, "qfe_batch_submit: request state exceeds the HBM budget");
    Request r;
    r.id = b->next_id++;
    r.prog = p;
    r.params.assign(params, params + np);
    b->queue.push_back(std::move(r));
    *request_id = b->queue.back().id;
    return QFE_OK;
}

int qfe_batch_step(qfe_batch b, uint64_t* done_ids, double* done_energies, uint64_t cap, uint64_t* ndone) {
    if (!b || !ndone || (cap && (!done_ids || !done_energies))) return fail(QFE_ERR_ARG, "qfe_batch_step: invalid argument");
    *ndone = 0;
    // admit
    while (!b->queue.empty() && (int)b->active.size() < b->max_active) {
        const uint64_t foot = 16ull << program_nqubit(b->queue.front().prog);
        if (b->used + foot > b->budget && !b->active.empty()) break;
        b->used += foot;
        b->active.push_back(std::move(b->queue.front()));
        b->queue.pop_front();
    }
    // advance: every active request runs its program; results retire in admission order
    std::vector<Request> still;
    for (auto& r : b->active) {
        if (*ndone >= cap) {
            still.push_back(std::move(r));
            continue;
        }
        double e = 0.0;
        int rc = qfe_program_energy(r.prog, r.params.data(), &e);
        if (rc) return rc;
        done_ids[*ndone] = r.id;
        done_energies[*ndone] = e;
        ++*ndone;
        b->used -= 16ull << program_nqubit(r.prog);
    }
    b->active.swap(still);
    return QFE_OK;
}

int qfe_batch_pending(qfe_batch b, uint64_t* queued, uint64_t* active) {
    if (!b) return fail(QFE_ERR_ARG, "qfe_batch_pending: invalid argument");
    if (queued) *queued = b->queue.size();
    if (active) *active = b->active.size();
    return QFE_OK;
}

}  // extern "C"
