// serve.hip — a request loop for MD segments with continuous batching.
//
// A segment is one independent system plus a step count. Submitted segments
// wait in FIFO order and are admitted while their device footprint fits the
// HBM budget; admitted segments stay resident until they finish. Every call to
// atvh_server_advance admits what fits, then advances all resident segments
// by one chunk (the smallest of the chunk length and every member's remaining
// steps, so nobody overshoots): every segment on the brick layout goes
// through one batched step (batch.hip; PairLJCut members share its pair
// launch, charged members queue their own pair and k-space calls inside it),
// the rest (small grids) through their own fused loop; finished
// segments are downloaded to host memory and their device memory is released,
// so later arrivals join at the next chunk. A member's trajectory does not
// depend on what it was batched with (batch invariance), nor on how its run
// was cut into chunks (the loop is continuable).
//
// The reference has no counterpart (its run() is continuable and frames move
// host<->device through to_cuda/to_cpu, include/atomiverse/cuda/dynamics/
// md.cuh:130-178, src/cuda/frame/transfer.cu:145-475); this is the serving
// layer over that boundary.

#include "atvh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <memory>
#include <string>
#include <vector>

using namespace atvh;

namespace {

struct Segment {
    int64_t id = 0;
    atvh_segment spec{};
    std::vector<double> x, y, z, vx, vy, vz, mass, charge, eps_tab, sigma_tab;
    std::vector<int> type, set_tab;
    int64_t bytes = 0;
    atvh_system* sys = nullptr;
    int64_t left = 0;
    bool batched = false;
    // results
    std::vector<double> rx, ry, rz, rvx, rvy, rvz;
    double pe = 0, ke = 0;
    // a segment whose admission or advance failed: the error it ended with
    int fail_code = 0;
    std::string fail_msg;
    Segment() = default;
    Segment(const Segment&) = delete;
    Segment& operator=(const Segment&) = delete;
    ~Segment() {
        if (sys)
            atvh_system_destroy(sys);
    }
};

template <class T>
std::vector<T> copy_or_empty(const T* p, int64_t n) {
    return p ? std::vector<T>(p, p + n) : std::vector<T>();
}

} // namespace

struct atvh_server {
    atvh_ctx* ctx = nullptr;
    int64_t budget = 0;
    int64_t in_use = 0;
    int chunk = 0;
    int precision = 0;
    int64_t next_id = 1;
    std::deque<std::unique_ptr<Segment>> queue;
    std::vector<std::unique_ptr<Segment>> resident;
    std::vector<std::unique_ptr<Segment>> done;
    std::vector<int64_t> newly_done;
    atvh_batch* batch = nullptr;
    std::vector<atvh_system*> batch_members;
};

namespace {

void check_rc(int rc, atvh_ctx* c, const char* what) {
    if (rc != ATVH_OK)
        throw Error(rc, std::string(what) + ": " + c->err);
}

// Device bytes a segment occupies: per-atom state and sort scratch
// (~45 words), the 16-bit full neighbour list at 1.5x the mean count + 32,
// per-brick halo tables, and for k-space segments the PME mesh (fixed-point,
// real and half spectrum grids of both precisions, influence table) plus two
// FFT plan pairs at the largest footprint one pair has taken in this process
// (64 MiB each before any plan was made).
int64_t estimate_bytes(const atvh_segment& g) {
    const double vol = g.box[0] * g.box[1] * g.box[2];
    const double rcs = g.cutoff + (g.skin < 0 ? 0.3 * g.cutoff : g.skin);
    const double mean = double(g.n) / vol * 4.18879020478639 * rcs * rcs * rcs;
    const double cap = std::ceil(1.5 * mean + 32) + 8;
    double bytes = double(g.n) * (45.0 * 8.0 + 2.0 * cap);
    const double cells = vol / std::pow(0.5 * rcs, 3);
    const double bricks = std::max(1.0, cells / 27.0);
    bytes += bricks * (2048.0 * 4 + 769.0 * 4 + 768.0 * 4 + 16);
    bytes += cells * 20.0;
    if (g.pair_style == 1 && g.pme_order > 0) {
        double K = 1;
        for (int d = 0; d < 3; ++d)
            K *= std::ceil(g.box[d] / g.pme_spacing) + 8;
        bytes += K * (8.0 + 8.0 + 4.0 + 8.0 + 4.0 + 4.0);
        bytes += 2.0 * static_cast<double>(fft_pair_bytes()); // FFT plan pairs of both precisions
    }
    return static_cast<int64_t>(bytes * 1.1 + (1 << 20));
}

void admit(atvh_server* sv, Segment& sg) {
    atvh_ctx* c = sv->ctx;
    const atvh_segment& g = sg.spec;
    check_rc(atvh_system_create(c, g.n, g.box, &sg.sys), c, "segment system");
    auto ptr = [](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };
    check_rc(atvh_frame_upload(sg.sys, sg.x.data(), sg.y.data(), sg.z.data(), ptr(sg.vx), ptr(sg.vy), ptr(sg.vz),
                               nullptr, nullptr, nullptr, sg.mass.data(), sg.type.empty() ? nullptr : sg.type.data(),
                               ptr(sg.charge)),
             c, "segment upload");
    check_rc(atvh_nlist_configure(sg.sys, g.cutoff, g.skin, 1, 0, 1), c, "segment neighbor list");
    if (g.pair_style == 1) {
        check_rc(atvh_pair_ljcoul_configure(sg.sys, g.ntypes, sg.eps_tab.data(), sg.sigma_tab.data(),
                                            sg.set_tab.data(), g.alpha, g.cutoff, 10000),
                 c, "segment pair style");
        if (g.pme_order > 0)
            check_rc(atvh_pme_configure(sg.sys, g.alpha, g.pme_spacing, g.pme_order), c, "segment k-space");
    } else {
        check_rc(atvh_pair_lj_configure(sg.sys, g.eps, g.sigma, g.cutoff, 0, nullptr, nullptr), c,
                 "segment pair style");
    }
    check_rc(atvh_integrator_configure(sg.sys, g.dt, g.thermostat, g.T, g.tau, 0), c, "segment integrator");
    check_rc(atvh_nlist_build(sg.sys), c, "segment list build");
    sg.batched = sg.sys->brick_mode;
    sg.left = g.nsteps;
}

void retire(atvh_server* sv, Segment& sg) {
    atvh_ctx* c = sv->ctx;
    const int64_t n = sg.spec.n;
    for (auto* v : {&sg.rx, &sg.ry, &sg.rz, &sg.rvx, &sg.rvy, &sg.rvz})
        v->assign(n, 0.0);
    check_rc(atvh_frame_download(sg.sys, sg.rx.data(), sg.ry.data(), sg.rz.data(), sg.rvx.data(), sg.rvy.data(),
                                 sg.rvz.data(), nullptr, nullptr, nullptr),
             c, "segment download");
    double vir[6];
    check_rc(atvh_energy(sg.sys, &sg.pe, vir, &sg.ke), c, "segment energy");
    atvh_system_destroy(sg.sys);
    sg.sys = nullptr;
    sv->in_use -= sg.bytes;
    // host copies of the inputs are no longer needed
    for (auto* v : {&sg.x, &sg.y, &sg.z, &sg.vx, &sg.vy, &sg.vz, &sg.mass, &sg.charge})
        std::vector<double>().swap(*v);
}

// A segment that failed is retired without results: its device memory is
// released and poll reports it like a finished one; fetch returns its error.
void fail_segment(atvh_server* sv, std::unique_ptr<Segment> sg, bool counted, int code, const std::string& msg) {
    if (sg->sys) {
        atvh_system_destroy(sg->sys);
        sg->sys = nullptr;
    }
    if (counted)
        sv->in_use -= sg->bytes;
    sg->fail_code = code != ATVH_OK ? code : ATVH_E_STATE;
    sg->fail_msg = msg;
    sv->newly_done.push_back(sg->id);
    sv->done.push_back(std::move(sg));
}

void drop_batch(atvh_server* sv) {
    if (sv->batch)
        atvh_batch_destroy(sv->batch);
    sv->batch = nullptr;
    sv->batch_members.clear();
}

} // namespace

#define API_TRY try {
#define API_CATCH(C)                                                                    \
    }                                                                                   \
    catch (const atvh::Error& e_) {                                                     \
        if (C)                                                                          \
            (C)->err = e_.what();                                                       \
        return e_.code;                                                                 \
    }                                                                                   \
    catch (const std::exception& e_) {                                                  \
        if (C)                                                                          \
            (C)->err = e_.what();                                                       \
        return ATVH_E_ARG;                                                              \
    }

extern "C" {

int64_t atvh_segment_bytes(const atvh_segment* seg) {
    if (!seg || seg->n <= 0)
        return ATVH_E_ARG;
    return estimate_bytes(*seg);
}

int atvh_server_create(atvh_ctx* ctx, int64_t hbm_budget, int chunk_steps, int precision, atvh_server** out) {
    if (!ctx || !out)
        return ATVH_E_ARG;
    API_TRY
    if (hbm_budget <= 0 || chunk_steps < 1)
        throw Error(ATVH_E_ARG, "atvh_server_create: budget and chunk must be positive");
    if (precision != ATVH_PREC_FP64 && precision != ATVH_PREC_MIXED)
        throw Error(ATVH_E_ARG, "atvh_server_create: unknown precision");
    auto* sv = new atvh_server();
    sv->ctx = ctx;
    sv->budget = hbm_budget;
    sv->chunk = chunk_steps;
    sv->precision = precision;
    ctx->live_systems += 1;
    *out = sv;
    return ATVH_OK;
    API_CATCH(ctx)
}

int atvh_server_destroy(atvh_server* sv) {
    if (!sv)
        return ATVH_E_ARG;
    drop_batch(sv);
    sv->resident.clear(); // each segment releases its device system
    sv->ctx->live_systems -= 1;
    delete sv;
    return ATVH_OK;
}

int atvh_server_submit(atvh_server* sv, const atvh_segment* g, int64_t* id) {
    if (!sv || !g)
        return ATVH_E_ARG;
    atvh_ctx* c = sv->ctx;
    API_TRY
    if (g->n <= 0 || !g->x || !g->y || !g->z || !g->mass || g->nsteps < 0)
        throw Error(ATVH_E_ARG, "atvh_server_submit: incomplete segment");
    if (g->pair_style == 1 && (g->ntypes < 1 || !g->eps_tab || !g->sigma_tab || !g->set_tab))
        throw Error(ATVH_E_ARG, "atvh_server_submit: the LJ + Coulomb style needs its coefficient table");
    auto sg = std::make_unique<Segment>();
    sg->id = sv->next_id++;
    sg->spec = *g;
    const int64_t n = g->n;
    sg->x = copy_or_empty(g->x, n);
    sg->y = copy_or_empty(g->y, n);
    sg->z = copy_or_empty(g->z, n);
    sg->vx = copy_or_empty(g->vx, n);
    sg->vy = copy_or_empty(g->vy, n);
    sg->vz = copy_or_empty(g->vz, n);
    sg->mass = copy_or_empty(g->mass, n);
    sg->charge = copy_or_empty(g->charge, n);
    sg->type = copy_or_empty(g->type, n);
    if (g->pair_style == 1) {
        const int n2 = g->ntypes * g->ntypes;
        sg->eps_tab = copy_or_empty(g->eps_tab, n2);
        sg->sigma_tab = copy_or_empty(g->sigma_tab, n2);
        sg->set_tab = copy_or_empty(g->set_tab, n2);
    }
    // the host pointers of the spec are not kept
    sg->spec.x = sg->spec.y = sg->spec.z = sg->spec.vx = sg->spec.vy = sg->spec.vz = nullptr;
    sg->spec.mass = sg->spec.charge = nullptr;
    sg->spec.type = nullptr;
    sg->spec.eps_tab = sg->spec.sigma_tab = nullptr;
    sg->spec.set_tab = nullptr;
    sg->bytes = estimate_bytes(*g);
    if (sg->bytes > sv->budget)
        throw Error(ATVH_E_CAPACITY, "atvh_server_submit: the segment alone exceeds the HBM budget");
    if (id)
        *id = sg->id;
    sv->queue.push_back(std::move(sg));
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_server_advance(atvh_server* sv, int* active, int* queued) {
    if (!sv)
        return ATVH_E_ARG;
    atvh_ctx* c = sv->ctx;
    API_TRY
    ATVH_HIP(hipSetDevice(c->device));
    // admission: FIFO while the head fits the budget
    bool changed = false;
    while (!sv->queue.empty() && sv->in_use + sv->queue.front()->bytes <= sv->budget) {
        std::unique_ptr<Segment> sg = std::move(sv->queue.front());
        sv->queue.pop_front();
        try {
            admit(sv, *sg);
        } catch (const Error& e) {
            fail_segment(sv, std::move(sg), false, e.code, e.what());
            continue;
        } catch (const std::exception& e) {
            fail_segment(sv, std::move(sg), false, ATVH_E_ARG, e.what());
            continue;
        }
        sv->in_use += sg->bytes;
        sv->resident.push_back(std::move(sg));
        changed = true;
    }
    if (!sv->resident.empty()) {
        int64_t step = sv->chunk;
        for (auto& r : sv->resident)
            step = std::min(step, r->left);
        // batched members: rebuild the batch when membership changed
        std::vector<atvh_system*> members;
        for (auto& r : sv->resident)
            if (r->batched && r->left > 0)
                members.push_back(r->sys);
        if (changed || members != sv->batch_members) {
            drop_batch(sv);
            if (!members.empty()) {
                check_rc(atvh_batch_create(c, members.data(), static_cast<int>(members.size()), &sv->batch), c,
                         "server batch");
                atvh_batch_set_graph(sv->batch, 1);
                sv->batch_members = members;
            }
        }
        if (step > 0) {
            // a run that fails ends its segments (marked failed, retired
            // below); only the segments whose run succeeded count the chunk
            if (sv->batch) {
                const int rc = atvh_batch_run(sv->batch, static_cast<int>(step), sv->precision, 0);
                if (rc != ATVH_OK) {
                    for (auto& r : sv->resident)
                        if (r->batched && r->left > 0) {
                            r->fail_code = rc;
                            r->fail_msg = std::string("server step: ") + c->err;
                        }
                    drop_batch(sv);
                }
            }
            for (auto& r : sv->resident)
                if (!r->batched && r->left > 0) {
                    const int rc = atvh_md_run(r->sys, static_cast<int>(step), sv->precision, 0);
                    if (rc != ATVH_OK) {
                        r->fail_code = rc;
                        r->fail_msg = std::string("server step: ") + c->err;
                    }
                }
            for (auto& r : sv->resident)
                if (r->fail_code == ATVH_OK)
                    r->left -= step;
        }
        // retire finished and failed segments (their device memory is released)
        std::vector<std::unique_ptr<Segment>> keep;
        for (auto& r : sv->resident) {
            if (r->fail_code != ATVH_OK) {
                const int code = r->fail_code;
                const std::string msg = r->fail_msg;
                fail_segment(sv, std::move(r), true, code, msg);
            } else if (r->left <= 0) {
                retire(sv, *r);
                sv->newly_done.push_back(r->id);
                sv->done.push_back(std::move(r));
            } else {
                keep.push_back(std::move(r));
            }
        }
        if (keep.size() != sv->resident.size())
            drop_batch(sv);
        sv->resident.swap(keep);
    }
    if (active)
        *active = static_cast<int>(sv->resident.size());
    if (queued)
        *queued = static_cast<int>(sv->queue.size());
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_server_poll(atvh_server* sv, int64_t* ids, int cap, int* n) {
    if (!sv || !n)
        return ATVH_E_ARG;
    const int k = std::min<int>(cap, static_cast<int>(sv->newly_done.size()));
    for (int i = 0; i < k; ++i)
        ids[i] = sv->newly_done[i];
    sv->newly_done.erase(sv->newly_done.begin(), sv->newly_done.begin() + k);
    *n = k;
    return ATVH_OK;
}

int atvh_server_fetch(atvh_server* sv, int64_t id, double* x, double* y, double* z, double* vx, double* vy,
                      double* vz, double* pe, double* ke) {
    if (!sv)
        return ATVH_E_ARG;
    atvh_ctx* c = sv->ctx;
    API_TRY
    auto it = std::find_if(sv->done.begin(), sv->done.end(), [&](const auto& s) { return s->id == id; });
    if (it == sv->done.end())
        throw Error(ATVH_E_STATE, "atvh_server_fetch: no finished segment with that id");
    Segment& sg = **it;
    if (sg.fail_code != ATVH_OK) {
        const int code = sg.fail_code;
        const std::string msg = "segment " + std::to_string(id) + " failed: " + sg.fail_msg;
        sv->done.erase(it);
        throw Error(code, msg);
    }
    auto put = [&](double* dst, const std::vector<double>& src) {
        if (dst)
            std::copy(src.begin(), src.end(), dst);
    };
    put(x, sg.rx);
    put(y, sg.ry);
    put(z, sg.rz);
    put(vx, sg.rvx);
    put(vy, sg.rvy);
    put(vz, sg.rvz);
    if (pe)
        *pe = sg.pe;
    if (ke)
        *ke = sg.ke;
    sv->done.erase(it);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_server_usage(atvh_server* sv, int64_t* bytes_in_use, int64_t* budget) {
    if (!sv)
        return ATVH_E_ARG;
    if (bytes_in_use)
        *bytes_in_use = sv->in_use;
    if (budget)
        *budget = sv->budget;
    return ATVH_OK;
}

} // extern "C"
