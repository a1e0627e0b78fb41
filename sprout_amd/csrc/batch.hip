// batch.hip — batches of tanks: tanks of the same grid and solver schedule advance together,
// every kernel of a substep launched once for a group of up to SP_LAUNCH_TANKS tanks
// (blockIdx.z = tank, schedule.hip), each group on a stream of its own so that the groups
// overlap on the device. A tank's results are those it gets alone: the same kernels on its
// own state, each tank's solves gated by its own convergence (a converged tank's blocks
// return at entry while the rest of its group iterates), its own sort, gather decision and
// iteration adaptation. The batch records as many iterations as its most demanding tank.
//
// Part of the libsprout_hip.so translation unit (included at the end of sprout.hip).
#include <memory>

struct sp_batch {
    struct Group {
        std::vector<sp_sim*> tanks;
        Run run;
        GraphSet gs;
        Timing tm;
        hipStream_t stream = nullptr;
        ~Group() {
            if (stream) (void)hipStreamSynchronize(stream);
            gs.clear();
            if (stream) (void)hipStreamDestroy(stream);
        }
    };
    int device = 0;
    std::vector<sp_sim*> sims;
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<uint64_t> keys;  // the grouping the groups were built for
    std::vector<uint64_t> gens;  // and the tanks' arena generations (an evicted and restored tank moved)
    uint32_t recorded = 0;       // substeps of the latest sp_batch_record (none: nothing to read back)
    std::string err;
};

namespace {

void set_batch_err(sp_batch* b, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    b->err = buf;
}

// Which group a tank belongs to: its grid and the schedule its tables ask for.
uint64_t group_key(const sp_sim* s) {
    return (uint64_t)(uint32_t)s->W | (uint64_t)(uint32_t)s->H << 16 | (uint64_t)schedule_key(s->gp) << 32;
}

// (Re)builds the groups when a tank's key changed or its arena moved: tanks in the order
// given, same key together, at most SP_LAUNCH_TANKS per group.
void regroup(sp_batch* b) {
    std::vector<uint64_t> keys, gens;
    for (sp_sim* s : b->sims) {
        keys.push_back(group_key(s));
        gens.push_back(s->generation);
    }
    if (keys == b->keys && gens == b->gens && !b->groups.empty()) return;
    b->groups.clear();
    b->keys = keys;
    b->gens = gens;
    std::vector<bool> taken(b->sims.size(), false);
    for (size_t i = 0; i < b->sims.size(); i++) {
        if (taken[i]) continue;
        std::unique_ptr<sp_batch::Group> g;
        for (size_t j = i; j < b->sims.size(); j++) {
            if (taken[j] || keys[j] != keys[i]) continue;
            if (!g) {
                g.reset(new sp_batch::Group());
                HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
                g->tm.create();
            }
            g->tanks.push_back(b->sims[j]);
            taken[j] = true;
            if ((int)g->tanks.size() == SP_LAUNCH_TANKS) {
                b->groups.push_back(std::move(g));
            }
        }
        if (g) b->groups.push_back(std::move(g));
    }
    for (auto& g : b->groups) {
        const sp_sim* f = g->tanks.front();
        Run& r = g->run;
        r = f->run;
        r.T = (int)g->tanks.size();
        for (int k = 0; k < r.T; k++) r.tb.t[k] = g->tanks[k]->ref;
        r.stream = g->stream;
        const char* gv = getenv("SPROUT_GRAPHS");
        g->gs.on = !(gv && gv[0] == '0');
    }
}

template <class F>
int batch_guarded(sp_batch* b, F&& f) {
    if (!b) return SP_E_INVALID;
    try {
        return f();
    } catch (const HipError& e) {
        set_batch_err(b, "%s failed: %s", e.what, hipGetErrorString(e.e));
        return SP_E_DEVICE;
    } catch (const std::bad_alloc&) {
        set_batch_err(b, "host allocation failed");
        return SP_E_NOMEM;
    } catch (...) {
        set_batch_err(b, "unexpected error");
        return SP_E_DEVICE;
    }
}

}  // namespace

extern "C" {

int sp_batch_create(sp_sim* const* sims, uint32_t n, sp_batch** out) {
    if (!sims || !out || n == 0) return SP_E_INVALID;
    *out = nullptr;
    for (uint32_t k = 0; k < n; k++) {
        if (!sims[k] || sims[k]->device != sims[0]->device) return SP_E_INVALID;
        for (uint32_t j = 0; j < k; j++)
            if (sims[j] == sims[k]) return SP_E_INVALID;
    }
    sp_batch* b = new (std::nothrow) sp_batch();
    if (!b) return SP_E_NOMEM;
    b->device = sims[0]->device;
    b->sims.assign(sims, sims + n);
    *out = b;
    return SP_OK;
}

void sp_batch_destroy(sp_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    b->groups.clear();
    delete b;
}

const char* sp_batch_last_error(const sp_batch* b) { return b ? b->err.c_str() : "null batch"; }

int sp_batch_record(sp_batch* b, uint32_t substeps, float target_s) {
    return batch_guarded(b, [&]() -> int {
        HIPCHK(hipSetDevice(b->device));
        b->recorded = substeps;
        for (size_t k = 0; k < b->sims.size(); k++) {
            sp_sim* s = b->sims[k];
            int rc = guarded(s, [&]() -> int { return record_prepare(s, substeps, target_s); });
            if (rc != SP_OK) {
                set_batch_err(b, "tank %zu: %s", k, s->err.c_str());
                return rc;
            }
        }
        // the tanks' uploads on their own streams finish before the groups' launches
        for (sp_sim* s : b->sims) HIPCHK(hipStreamSynchronize(s->stream));
        regroup(b);
        for (auto& g : b->groups) {
            Run& r = g->run;
            r.slots = 0;
            for (sp_sim* s : g->tanks) r.slots = std::max(r.slots, s->n_slots);
        }
        for (uint32_t k = 0; k < substeps; k++) {
            for (auto& g : b->groups) {
                Recs rec{0, 0, 0};
                for (sp_sim* s : g->tanks) {
                    rec.p = std::max(rec.p, s->rec.p);
                    rec.p2 = std::max(rec.p2, s->rec.p2);
                    rec.v = std::max(rec.v, s->rec.v);
                    maybe_sort(s, g->stream);
                }
                launch_substep(g->run, &g->gs, &g->tm, -1, rec);
                for (sp_sim* s : g->tanks) s->grid_valid = true;
            }
        }
        return SP_OK;
    });
}

int sp_batch_wait(sp_batch* b) {
    return batch_guarded(b, [&]() -> int {
        HIPCHK(hipSetDevice(b->device));
        for (auto& g : b->groups) {
            HIPCHK(hipStreamSynchronize(g->stream));
            g->tm.collect();
        }
        for (size_t k = 0; k < b->sims.size(); k++) {
            sp_sim* s = b->sims[k];
            int rc = guarded(s, [&]() -> int {
                // a record of no substeps reads nothing back (record, mod.rs:2559-2561)
                if (b->recorded > 0) readback(s);
                else poll_readbacks(s, true);
                return SP_OK;
            });
            if (rc != SP_OK) {
                set_batch_err(b, "tank %zu: %s", k, s->err.c_str());
                return rc;
            }
        }
        return SP_OK;
    });
}

int sp_batch_groups(const sp_batch* b, uint32_t* n_groups, uint32_t* tanks_per_group, uint32_t cap) {
    if (!b || !n_groups) return SP_E_INVALID;
    *n_groups = (uint32_t)b->groups.size();
    if (tanks_per_group)
        for (uint32_t k = 0; k < cap && k < b->groups.size(); k++) tanks_per_group[k] = (uint32_t)b->groups[k]->tanks.size();
    return SP_OK;
}

int sp_batch_timings(const sp_batch* b, uint32_t group, float* out8) {
    if (!b || !out8 || group >= b->groups.size()) return SP_E_INVALID;
    const Timing& t = b->groups[group]->tm;
    for (int k = 0; k < 7; k++) out8[k] = t.phase[k];
    out8[7] = t.particle_ms;
    return SP_OK;
}

}  // extern "C"
