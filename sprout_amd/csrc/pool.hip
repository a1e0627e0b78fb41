// pool.hip — continuous batching of step requests over many resident tanks, driven by the
// caller's serving loop (sp_pool_*). The rounds, queues and admission are pool_core.h; this
// file is its device side:
//   - a request starts with the words record_prepare writes (the batch words of the step
//     state and the tank's recorded-iteration caps, mod.rs:2521-2526); k_pool_begin writes
//     them from its kernel arguments on the launch stream, ahead of the request's first
//     substep. A tank with spawns waiting or tables to upload takes record_prepare itself,
//     on its own stream, synchronised;
//   - a round's group is one substep of launch_substep over its tanks (blockIdx.z), on one
//     of the pool's streams, with the tanks' sorts first (as sp_batch_record does); launches
//     are direct, not graphs, since a group's tanks change from round to round;
//   - after the last substep of a request, k_pool_finish copies the tank's step and solve
//     records into the launch's mailbox and clears the solve maxima (what sp_wait's readback
//     reads and clears), and one copy brings the mailbox to pinned host memory;
//   - the round ends when every stream is synchronised (as sp_batch_wait); the finished
//     requests' readback is then folded in from the pinned copies, with no device call;
//   - evictions and restores run between rounds, with no launch in flight (as sp_evict and
//     the restore of a later call).
// No thread of its own and no event polling: the thread that calls sp_pool_step,
// sp_pool_wait or sp_pool_drain runs the round on the device.
//
// Part of the libsprout_hip.so translation unit (included at the end of sprout.hip).
#include <memory>
#include <stdexcept>

#include "pool_core.h"

// Kernel arguments of a launch's request starts: for tank t of the launch (bit t of `mask`),
// the step words and caps record_prepare would write.
struct PoolStart {
    uint32_t mask;
    float target[SP_LAUNCH_TANKS];
    uint32_t substeps[SP_LAUNCH_TANKS];
    uint32_t caps[SP_LAUNCH_TANKS][3];
};

// A tank's records as sp_wait reads them back.
struct PoolRec {
    sp_stepbuf S;
    sp_solve K;
};

__global__ __launch_bounds__(64) void k_pool_begin(const TankBatch tb, const PoolStart ps) {
    const int t = blockIdx.z;
    if (threadIdx.x != 0 || !((ps.mask >> t) & 1u)) return;
    const Ctx c = ctx_of(tb.t[t]);
    c.S->frame_left = ps.target[t];
    c.S->batch_steps = ps.substeps[t];
    c.S->batch_done = 0u;
    c.K->cap_p = ps.caps[t][0];
    c.K->cap_p2 = ps.caps[t][1];
    c.K->cap_v = ps.caps[t][2];
}

__global__ __launch_bounds__(64) void k_pool_finish(const TankBatch tb, const uint32_t mask, PoolRec* out) {
    const int t = blockIdx.z;
    if (!((mask >> t) & 1u)) return;
    const Ctx c = ctx_of(tb.t[t]);
    const uint32_t* S = reinterpret_cast<const uint32_t*>(c.S);
    const uint32_t* K = reinterpret_cast<const uint32_t*>(c.K);
    uint32_t* oS = reinterpret_cast<uint32_t*>(&out[t].S);
    uint32_t* oK = reinterpret_cast<uint32_t*>(&out[t].K);
    for (uint32_t i = threadIdx.x; i < sizeof(sp_stepbuf) / 4; i += 64) oS[i] = S[i];
    for (uint32_t i = threadIdx.x; i < sizeof(sp_solve) / 4; i += 64) oK[i] = K[i];
    __syncthreads();
    if (threadIdx.x == 0) {  // clear_maxima
        c.K->it_max = 0u;
        c.K->p2_it_max = 0u;
        c.K->v_it_max = 0u;
    }
}

static_assert(sizeof(sp_stepbuf) % 4 == 0 && sizeof(sp_solve) % 4 == 0, "records copied as words");
static_assert(sizeof(PoolStart) <= 2048, "kernel argument size");

struct PoolBackend {
    int device = 0;
    std::vector<sp_sim*> sims;
    std::vector<hipStream_t> streams;
    std::vector<char> used;  // streams with launches in this round
    std::vector<std::unique_ptr<Timing>> tms;  // per stream: timings for the launch-shape tuner
    struct Start {
        bool pending = false;  // written by the first launch of the request
        float target = 0.0f;
        uint32_t substeps = 0;
        uint32_t caps[3] = {0, 0, 0};
    };
    std::vector<Start> starts;  // per scene
    struct Mailbox {
        PoolRec* dev = nullptr;
        PoolRec* host = nullptr;  // pinned
    };
    std::vector<Mailbox> boxes;  // per launch of a round
    size_t n_launch = 0;
    std::vector<std::pair<uint32_t, uint32_t>> where;  // per scene: (launch, tank) of its records

    Mailbox& box(size_t i) {
        while (boxes.size() <= i) {
            Mailbox b;
            HIPCHK(hipMalloc((void**)&b.dev, sizeof(PoolRec) * SP_LAUNCH_TANKS));
            hipError_t e = hipHostMalloc((void**)&b.host, sizeof(PoolRec) * SP_LAUNCH_TANKS, hipHostMallocDefault);
            if (e != hipSuccess) {
                (void)hipFree(b.dev);
                throw HipError{e, "hipHostMalloc (pool mailbox)"};
            }
            boxes.push_back(b);
        }
        return boxes[i];
    }

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
        Start& st = starts[k];
        if (!s->params_set || !s->pending_spawn.empty() || s->tables_dirty) {
            const int rc = record_prepare(s, substeps, target);
            if (rc != SP_OK) return rc;
            HIPCHK(hipStreamSynchronize(s->stream));
            st.pending = false;
            return SP_OK;
        }
        st.pending = true;
        st.target = std::max(target, 0.0f);
        st.substeps = substeps;
        st.caps[0] = s->rec.p;
        st.caps[1] = s->rec.p2;
        st.caps[2] = s->rec.v;
        return SP_OK;
    }

    void launch(int si, const std::vector<uint32_t>& members, const std::vector<char>& last) {
        if (members.empty() || members.size() > SP_LAUNCH_TANKS || last.size() != members.size())
            throw std::runtime_error("pool launch group out of range");
        const hipStream_t stream = streams[(size_t)si];
        sp_sim* f = sims[members[0]];
        Run r = f->run;
        r.T = (int)members.size();
        r.stream = stream;
        r.slots = 0;
        Recs rec{0, 0, 0};
        PoolStart ps{};
        uint32_t last_mask = 0;
        for (int k = 0; k < r.T; k++) {
            const uint32_t m = members[(size_t)k];
            sp_sim* s = sims[m];
            r.tb.t[k] = s->ref;
            r.slots = std::max(r.slots, s->n_slots);
            rec.p = std::max(rec.p, s->rec.p);
            rec.p2 = std::max(rec.p2, s->rec.p2);
            rec.v = std::max(rec.v, s->rec.v);
            Start& st = starts[m];
            if (st.pending) {
                ps.mask |= 1u << k;
                ps.target[k] = st.target;
                ps.substeps[k] = st.substeps;
                for (int j = 0; j < 3; j++) ps.caps[k][j] = st.caps[j];
                st.pending = false;
            }
            if (last[(size_t)k]) last_mask |= 1u << k;
        }
        const dim3 per_tank(1, 1, (unsigned)r.T);
        if (ps.mask) hipLaunchKernelGGL(k_pool_begin, per_tank, dim3(64), 0, stream, r.tb, ps);
        for (uint32_t m : members) maybe_sort(sims[m], stream);
        Timing* tm = tuner().sampling(particle_key(r)) ? tms[(size_t)si].get() : nullptr;
        launch_substep(r, nullptr, tm, -1, rec);
        if (last_mask) {
            Mailbox& b = box(n_launch);
            hipLaunchKernelGGL(k_pool_finish, per_tank, dim3(64), 0, stream, r.tb, last_mask, b.dev);
            HIPCHK(hipMemcpyAsync(b.host, b.dev, sizeof(PoolRec) * (size_t)r.T, hipMemcpyDeviceToHost, stream));
            for (int k = 0; k < r.T; k++)
                if (last[(size_t)k]) where[members[(size_t)k]] = {(uint32_t)n_launch, (uint32_t)k};
        }
        HIPCHK(hipGetLastError());
        for (uint32_t m : members) sims[m]->grid_valid = true;
        used[(size_t)si] = 1;
        n_launch++;
    }

    void sync() {
        for (size_t i = 0; i < streams.size(); i++)
            if (used[i]) {
                HIPCHK(hipStreamSynchronize(streams[i]));
                tms[i]->collect();
                used[i] = 0;
            }
        n_launch = 0;
    }

    // The request's readback (sp_wait's, mod.rs:3033-3116) from the pinned copies.
    uint64_t finish(uint32_t k, uint32_t substeps) {
        sp_sim* s = sims[k];
        const auto w = where[k];
        const PoolRec& rec = boxes[w.first].host[w.second];
        readback_apply(s, rec.S, rec.K);
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
    be.tms.clear();
    for (auto& b : be.boxes) {
        (void)hipFree(b.dev);
        (void)hipHostFree(b.host);
    }
    be.boxes.clear();
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
        for (uint32_t i = 0; i < streams; i++) {
            be.tms.emplace_back(new Timing());
            be.tms.back()->create();
        }
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
    for (uint32_t k = 0; k < n; k++) v->be.sims[k]->pooled = false;
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
    if (s->device != v->be.device || s->pooled || s->in_batch || !s->params_set) return SP_E_INVALID;
    // the tables the launch groups are keyed on, while the handle is still the caller's, and
    // the read-backs of the batches it recorded itself
    int rc = guarded(s, [&]() -> int {
        HIPCHK(hipStreamSynchronize(s->stream));
        poll_readbacks(s, true);
        return upload_tables(s);
    });
    if (rc != SP_OK) return rc;
    if (big_bytes(s) > v->core->budget()) return SP_E_CAPACITY;
    bool ok = v->core->add(scene, [&](uint32_t id) {
        v->be.sims.push_back(s);
        v->be.starts.emplace_back();
        v->be.where.emplace_back(0u, 0u);
        (void)id;
        s->pooled = true;
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

int sp_pool_stats_get(sp_pool* v, sp_pool_stats* out) {
    if (!v || !out) return SP_E_INVALID;
    const pool::Stats st = v->core->stats();
    *out = sp_pool_stats{};
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
