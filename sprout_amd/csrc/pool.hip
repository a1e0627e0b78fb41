// pool.hip — continuous batching of step requests over many resident tanks, driven by the
// caller's serving loop (sp_pool_*). The rounds, queues and admission are pool_core.h; this
// file is its device side, made of the sequences the rest of the library already runs:
//   - a request starts with the tank's record_prepare on its own stream, then that stream
//     is synchronised (as sp_batch_record does for each tank);
//   - a round's group is one substep of launch_substep over its tanks (blockIdx.z), on one
//     of the pool's streams, with the tanks' sorts first (as sp_batch_record does); launches
//     are direct, not graphs, since a group's tanks change from round to round;
//   - the round ends when every stream is synchronised (as sp_batch_wait);
//   - a finished request's readback is the tank's own (as sp_wait);
//   - evictions and restores run between rounds, with no launch in flight (as sp_evict and
//     the restore of a later call).
// No thread of its own and no event polling: the thread that calls sp_pool_step,
// sp_pool_wait or sp_pool_drain runs the round on the device.
//
// Part of the libsprout_hip.so translation unit (included at the end of sprout.hip).
#include <memory>
#include <stdexcept>

#include "pool_core.h"

struct PoolBackend {
    int device = 0;
    std::vector<sp_sim*> sims;
    std::vector<hipStream_t> streams;
    std::vector<char> used;  // streams with launches in this round

    uint64_t bytes(uint32_t k) { return big_bytes(sims[k]); }
    bool resident(uint32_t k) { return sims[k]->resident; }
    void evict(uint32_t k) { ::evict(sims[k]); }
    void restore(uint32_t k) {
        HIPCHK(hipSetDevice(device));
        ::restore(sims[k]);
    }
    uint64_t key(uint32_t k) { return group_key(sims[k]); }

    int begin(uint32_t k, uint32_t substeps, float target) {
        HIPCHK(hipSetDevice(device));
        sp_sim* s = sims[k];
        const int rc = record_prepare(s, substeps, target);
        if (rc != SP_OK) return rc;
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    }

    void launch(int si, const std::vector<uint32_t>& members) {
        const hipStream_t stream = streams[(size_t)si];
        sp_sim* f = sims[members[0]];
        Run r = f->run;
        r.T = (int)members.size();
        r.stream = stream;
        r.slots = 0;
        Recs rec{0, 0, 0};
        for (int k = 0; k < r.T; k++) {
            sp_sim* s = sims[members[(size_t)k]];
            r.tb.t[k] = s->ref;
            r.slots = std::max(r.slots, s->n_slots);
            rec.p = std::max(rec.p, s->rec.p);
            rec.p2 = std::max(rec.p2, s->rec.p2);
            rec.v = std::max(rec.v, s->rec.v);
            maybe_sort(s, stream);
        }
        launch_substep(r, nullptr, nullptr, -1, rec);
        for (uint32_t m : members) sims[m]->grid_valid = true;
        used[(size_t)si] = 1;
    }

    void sync() {
        for (size_t i = 0; i < streams.size(); i++)
            if (used[i]) {
                HIPCHK(hipStreamSynchronize(streams[i]));
                used[i] = 0;
            }
    }

    uint64_t finish(uint32_t k, uint32_t substeps) {
        sp_sim* s = sims[k];
        readback(s);
        return (uint64_t)s->n_slots * substeps;
    }
};

struct sp_pool {
    PoolBackend be;
    std::unique_ptr<pool::Core<PoolBackend>> core;
    std::string err;
};

namespace {

void fill_pool_info(const pool::Req& q, sp_request_info* out) {
    out->ticket = q.ticket;
    out->scene = q.scene;
    out->substeps = q.substeps;
    out->status = q.status == pool::REQ_PENDING ? SP_E_BUSY : q.status;
    out->tanks = q.tanks;
    out->submit_ms = q.submit_ms;
    out->launch_ms = q.start_ms;
    out->done_ms = q.done_ms;
    out->latency_ms = q.status == pool::REQ_PENDING ? 0.0 : q.done_ms - q.submit_ms;
}

void free_pool_streams(PoolBackend& be) {
    for (hipStream_t st : be.streams)
        if (st) (void)hipStreamDestroy(st);
    be.streams.clear();
}

}  // namespace

extern "C" {

int sp_pool_create(int32_t device, uint64_t hbm_budget, uint32_t streams, sp_pool** out) {
    if (!out || streams == 0 || streams > 16 || hbm_budget == 0) return SP_E_INVALID;
    *out = nullptr;
    sp_pool* v = new (std::nothrow) sp_pool();
    if (!v) return SP_E_NOMEM;
    PoolBackend& be = v->be;
    be.device = device;
    try {
        HIPCHK(hipSetDevice(device));
        be.streams.assign(streams, nullptr);
        be.used.assign(streams, 0);
        for (auto& st : be.streams) HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    } catch (const HipError&) {
        free_pool_streams(be);
        delete v;
        return SP_E_DEVICE;
    }
    v->core.reset(new pool::Core<PoolBackend>(be, hbm_budget, SP_LAUNCH_TANKS, (int)streams, SP_E_DEVICE));
    *out = v;
    return SP_OK;
}

void sp_pool_destroy(sp_pool* v) {
    if (!v) return;
    (void)v->core->drain();
    const uint32_t n = v->core->scenes_added();
    for (uint32_t k = 0; k < n; k++) v->be.sims[k]->served = false;
    (void)hipSetDevice(v->be.device);
    free_pool_streams(v->be);
    delete v;
}

const char* sp_pool_last_error(sp_pool* v) {
    if (!v) return "null pool";
    v->err = v->core->error();
    return v->err.c_str();
}

int sp_pool_add(sp_pool* v, sp_sim* s, uint32_t* scene) {
    if (!v || !s || !scene) return SP_E_INVALID;
    if (s->device != v->be.device || s->served || s->in_batch || !s->params_set) return SP_E_INVALID;
    // the tables the launch groups are keyed on, while the handle is still the caller's
    int rc = guarded(s, [&]() -> int { return upload_tables(s); });
    if (rc != SP_OK) return rc;
    if (big_bytes(s) > v->core->budget()) return SP_E_CAPACITY;
    bool ok = v->core->add(scene, [&](uint32_t id) {
        v->be.sims.push_back(s);
        (void)id;
        s->served = true;
        return true;
    });
    return ok ? SP_OK : SP_E_DEVICE;
}

int sp_pool_submit(sp_pool* v, uint32_t scene, uint32_t substeps, float target_s, uint64_t* ticket) {
    if (!v || !ticket || substeps == 0 || !(target_s >= 0.0f)) return SP_E_INVALID;
    int rc = v->core->submit(scene, substeps, target_s, ticket);
    return rc == 0 ? SP_OK : rc == -1 ? SP_E_INVALID : SP_E_DEVICE;
}

int sp_pool_step(sp_pool* v, double wait_ms, uint32_t* advanced) {
    if (!v || !(wait_ms >= 0.0)) return SP_E_INVALID;
    const int n = v->core->step(wait_ms);
    if (advanced) *advanced = n > 0 ? (uint32_t)n : 0u;
    return n < 0 ? SP_E_DEVICE : SP_OK;
}

int sp_pool_poll(sp_pool* v, uint64_t ticket, sp_request_info* out) {
    if (!v || !out) return SP_E_INVALID;
    pool::Req q;
    if (!v->core->get(ticket, false, &q)) return SP_E_INVALID;
    fill_pool_info(q, out);
    return q.status == pool::REQ_PENDING ? SP_E_BUSY : SP_OK;
}

int sp_pool_wait(sp_pool* v, uint64_t ticket, sp_request_info* out) {
    if (!v || !out) return SP_E_INVALID;
    pool::Req q;
    if (!v->core->get(ticket, true, &q)) return SP_E_INVALID;
    fill_pool_info(q, out);
    return SP_OK;
}

int sp_pool_drain(sp_pool* v) {
    if (!v) return SP_E_INVALID;
    return v->core->drain() ? SP_OK : SP_E_DEVICE;
}

int sp_pool_stats_get(sp_pool* v, sp_server_stats* out) {
    if (!v || !out) return SP_E_INVALID;
    const pool::Stats st = v->core->stats();
    *out = sp_server_stats{};
    out->submitted = st.submitted;
    out->completed = st.completed;
    out->failed = st.failed;
    out->particle_steps = st.particle_steps;
    out->launches = st.launches;
    out->evictions = st.evictions;
    out->restores = st.restores;
    out->resident_bytes = st.resident_bytes;
    out->scenes = st.scenes;
    out->resident = st.resident;
    out->p50_ms = st.p50_ms;
    out->p99_ms = st.p99_ms;
    out->mean_ms = st.mean_ms;
    out->max_ms = st.max_ms;
    out->span_s = st.span_s;
    out->tanks_per_launch = st.tanks_per_launch;
    out->rounds = st.rounds;
    return SP_OK;
}

int sp_pool_reset_stats(sp_pool* v) {
    if (!v) return SP_E_INVALID;
    v->core->reset_stats();
    return SP_OK;
}

}  // extern "C"
