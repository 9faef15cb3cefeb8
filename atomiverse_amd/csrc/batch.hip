// batch.hip — many independent systems advanced by one launch per stage.
//
// The step of every member is the single-system step (integrate.hip's fused
// loop) with each stage launched once over all members: a member owns a
// contiguous range of the launch's blocks and runs exactly its own
// single-system code on its own buffers (batch.hpp), so a member's trajectory
// is bit-identical alone and in any batch. The per-step sequence — kick/drift,
// NH + decide, the ten predicated rebuild stages, pair forces, energy totals,
// final kick, NH + KE — is 16 or 17 launches whatever the number of members,
// and can be captured once into a hipGraph and replayed (no memset nodes: the
// step clears its counters with kernels).
//
// Reference counterpart: replicas advanced one after another
// (Remd::propagate_, src/cpu/path/remd.cpp:98-103), each a cpu::md::MD.

#include "atvh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace atvh;

namespace {

// A member's thermo record cleared to NaN (no energy asked at that step).
__global__ void k_fill_nan(double* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = __builtin_nan("");
}

// Which members a rebuild overflow halted (read back in one copy).
__global__ void k_batch_halted(DevCtl* const* ctls, int nsys, int* halted_out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nsys)
        halted_out[k] = halted(ctls[k]) ? 1 : 0;
}

__global__ void k_batch_reset(DevCtl** ctls, int nsys) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nsys) {
        ctls[k]->rec_step = 0;
        ctls[k]->steps_done = 0;
    }
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// One exchange round (Remd::exchange_round_, src/cpu/path/remd.cpp:105-151)
// between the members holding adjacent ladder slots (k, k+1), k = parity,
// parity + 2, ...: accept with min(1, exp((beta_k - beta_k+1)(U_k - U_k+1)))
// against a counter-based uniform u(seed, round, k). The reference swaps
// configurations and rescales velocities so each slot keeps its temperature;
// here the two members swap their thermostat (T, Q, Ndof kB T, xi) instead —
// the same exchange with configurations staying put — and each member's
// velocities are rescaled by sqrt(T_new / T_old) through its pending NH scale
// (the scale is applied in the next kick; the kinetic energy is rescaled now).
__global__ void k_batch_exchange(DevCtl** ctls, int* slot_member, int nslots, int parity,
                                 unsigned long long seed, unsigned long long round, int* accepted) {
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    int acc = 0;
    for (int k = parity; k + 1 < nslots; k += 2) {
        DevCtl* ci = ctls[slot_member[k]];
        DevCtl* cj = ctls[slot_member[k + 1]];
        const double bi = 1.0 / (kB_ev_per_k * ci->T_target);
        const double bj = 1.0 / (kB_ev_per_k * cj->T_target);
        const double arg = (bi - bj) * (ci->pe - cj->pe);
        const double p = arg >= 0.0 ? 1.0 : exp(arg);
        const unsigned long long r = splitmix64(seed ^ splitmix64((round << 32) ^ static_cast<unsigned long long>(k)));
        const double u = static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0);
        if (u < p) {
            ++acc;
            const double Ti = ci->T_target, Tj = cj->T_target;
            const double si = sqrt(Tj / Ti), sj = sqrt(Ti / Tj);
            ci->nh_scale *= si;
            ci->ke *= si * si;
            cj->nh_scale *= sj;
            cj->ke *= sj * sj;
            double t;
            t = ci->T_target; ci->T_target = cj->T_target; cj->T_target = t;
            t = ci->nh_Q; ci->nh_Q = cj->nh_Q; cj->nh_Q = t;
            t = ci->nh_ndof_kT; ci->nh_ndof_kT = cj->nh_ndof_kT; cj->nh_ndof_kT = t;
            t = ci->xi; ci->xi = cj->xi; cj->xi = t;
            const int m = slot_member[k];
            slot_member[k] = slot_member[k + 1];
            slot_member[k + 1] = m;
        }
    }
    *accepted = acc;
}

void free_graphs(atvh_batch* b) {
    for (int k = 0; k < 2; ++k) {
        if (b->graph[k])
            (void)hipGraphExecDestroy(b->graph[k]);
        b->graph[k] = nullptr;
        b->graph_prec[k] = -1;
    }
}

// One batched step. With timing on (an eager run: timed steps are never
// captured) each launch family is bracketed by events of its own, named
// apart from the single-system families (a charged member's own force call
// inside the pair family times itself as "pair").
void enqueue_step(atvh_batch* b, int precision, bool ef) {
    atvh_ctx* c = b->ctx;
    {
        ScopedTimer tm(c, "batch_integrate");
        integ_batch_launch(b, "initial");
        integ_batch_launch(b, "post_initial");
    }
    {
        ScopedTimer tm(c, "batch_nlist");
        nl_batch_launch_rebuild(b);
    }
    {
        ScopedTimer tm(c, "batch_pair");
        pair_batch_launch(b, precision, ef);
    }
    if (ef) {
        ScopedTimer tm(c, "batch_energy");
        integ_batch_launch(b, "finalize");
    }
    ScopedTimer tm(c, "batch_integrate");
    integ_batch_launch(b, "final");
    integ_batch_launch(b, "post_final");
}

void build_stages(atvh_batch* b, int nsteps) {
    // thermo buffers: one record per step per member, NaN where no energy was asked
    for (atvh_system* s : b->sys) {
        if (s->thermo_cap < nsteps) {
            dev_free(s->d_thermo);
            s->thermo_cap = nsteps;
            dev_alloc(&s->d_thermo, size_t(2) * nsteps);
        }
        s->thermo_n = nsteps;
        const int64_t nrec = int64_t(2) * s->thermo_cap;
        k_fill_nan<<<static_cast<int>(std::min<int64_t>((nrec + kBlock - 1) / kBlock, 1024)), kBlock, 0, b->ctx->stream>>>(
            s->d_thermo, nrec);
        ATVH_HIP(hipGetLastError());
    }
    batch_free_stages(b);
    free_graphs(b);
    integ_batch_build(b);
    nl_batch_build(b);
    pair_batch_build(b);
}

} // namespace

namespace atvh {

void batch_free_stages(atvh_batch* b) {
    for (auto& kv : b->stages) {
        dev_free(kv.second.args);
        dev_free(kv.second.base);
        dev_free(kv.second.owner);
    }
    b->stages.clear();
    dev_free(b->nl_nb);
    dev_free(b->nl_act);
    dev_free(b->nl_flags);
    dev_free(b->nl_next);
}

} // namespace atvh

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

int atvh_batch_create(atvh_ctx* ctx, atvh_system* const* systems, int nsys, atvh_batch** out) {
    if (!ctx || !out || !systems || nsys < 1)
        return ATVH_E_ARG;
    API_TRY
    ATVH_HIP(hipSetDevice(ctx->device));
    auto* b = new atvh_batch();
    b->ctx = ctx;
    try {
        std::vector<DevCtl*> ctls;
        for (int k = 0; k < nsys; ++k) {
            atvh_system* s = systems[k];
            if (!s || s->ctx != ctx)
                throw Error(ATVH_E_ARG, "atvh_batch_create: members must be systems of this context");
            for (int j = 0; j < k; ++j)
                if (systems[j] == s)
                    throw Error(ATVH_E_ARG, "atvh_batch_create: a system appears twice");
            b->sys.push_back(s);
            ctls.push_back(s->ctl);
        }
        dev_alloc(&b->d_ctls, ctls.size());
        ATVH_HIP(hipMemcpy(b->d_ctls, ctls.data(), ctls.size() * sizeof(DevCtl*), hipMemcpyHostToDevice));
        std::vector<int> slots(nsys);
        for (int k = 0; k < nsys; ++k)
            slots[k] = k;
        dev_alloc(&b->d_slot_member, slots.size());
        ATVH_HIP(hipMemcpy(b->d_slot_member, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice));
        dev_alloc(&b->d_accept, 1);
        dev_alloc(&b->d_halted, ctls.size());
    } catch (...) {
        dev_free(b->d_ctls);
        dev_free(b->d_slot_member);
        dev_free(b->d_accept);
        dev_free(b->d_halted);
        delete b;
        throw;
    }
    ctx->live_systems += 1; // a batch pins its context like a system does
    *out = b;
    return ATVH_OK;
    API_CATCH(ctx)
}

int atvh_batch_destroy(atvh_batch* b) {
    if (!b)
        return ATVH_E_ARG;
    atvh_ctx* c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_graphs(b);
    batch_free_stages(b);
    dev_free(b->d_ctls);
    dev_free(b->d_slot_member);
    dev_free(b->d_accept);
    dev_free(b->d_halted);
    c->live_systems -= 1;
    delete b;
    return ATVH_OK;
}

int atvh_batch_set_graph(atvh_batch* b, int enable) {
    if (!b)
        return ATVH_E_ARG;
    b->use_graph = enable != 0;
    free_graphs(b);
    return ATVH_OK;
}

int atvh_batch_run(atvh_batch* b, int nsteps, int precision, int energy_every) {
    if (!b)
        return ATVH_E_ARG;
    atvh_ctx* c = b->ctx;
    API_TRY
    if (nsteps < 0)
        throw Error(ATVH_E_ARG, "atvh_batch_run: nsteps < 0");
    if (precision != ATVH_PREC_FP64 && precision != ATVH_PREC_MIXED)
        throw Error(ATVH_E_ARG, "atvh_batch_run: unknown precision");
    ATVH_HIP(hipSetDevice(c->device));
    for (atvh_system* s : b->sys) {
        if (!s->integ_configured || !s->nl_configured || !s->pair_configured)
            throw Error(ATVH_E_STATE, "atvh_batch_run: every member needs an integrator, a neighbor list and "
                                      "a pair style");
        if (s->npt)
            throw Error(ATVH_E_STATE, "atvh_batch_run: NoseHooverNPT members run alone (atvh_md_run)");
        if (!s->nl_built) {
            const int rc = atvh_nlist_build(s);
            if (rc != ATVH_OK)
                throw Error(rc, std::string("atvh_batch_run: member list build failed: ") + c->err);
        }
    }
    if (nsteps == 0)
        return ATVH_OK;
    hipStream_t st = c->stream;
    // k-space members: every buffer and transform plan made (and each plan
    // run once) before the step is queued or captured
    for (atvh_system* s : b->sys)
        pme_prepare(s, precision);
    build_stages(b, nsteps);
    const int nsys = static_cast<int>(b->sys.size());
    k_batch_reset<<<(nsys + kBlock - 1) / kBlock, kBlock, 0, st>>>(b->d_ctls, nsys);
    ATVH_HIP(hipGetLastError());
    const bool timing = c->timing;
    // Steps that queue k-space members' transforms are not captured: the FFT
    // library's calls run eagerly between the batched launches (a serving run
    // that captured them faulted the GPU; the cause was not found)
    bool kspace = false;
    for (atvh_system* s : b->solo)
        kspace = kspace || s->pme != nullptr;
    for (int step = 0; step < nsteps; ++step) {
        const bool ef = (step == nsteps - 1) || (energy_every > 0 && (step % energy_every) == 0);
        if (b->use_graph && !timing && !kspace) {
            hipGraphExec_t& ge = b->graph[ef ? 1 : 0];
            if (!ge || b->graph_prec[ef ? 1 : 0] != precision) {
                if (ge)
                    ATVH_HIP(hipGraphExecDestroy(ge));
                ge = nullptr;
                hipGraph_t g = nullptr;
                ATVH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_step(b, precision, ef);
                } catch (...) {
                    (void)hipStreamEndCapture(st, &g);
                    if (g)
                        (void)hipGraphDestroy(g);
                    throw;
                }
                ATVH_HIP(hipStreamEndCapture(st, &g));
                const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                ATVH_HIP(e);
                b->graph_prec[ef ? 1 : 0] = precision;
            }
            ATVH_HIP(hipGraphLaunch(ge, st));
        } else {
            ScopedTimer tm(c, "batch_step");
            enqueue_step(b, precision, ef);
        }
    }
    // a member whose rebuild overflowed its list capacity or a brick halo
    // halted at that step; it regrows its list and finishes the run alone
    // (bit-identical to the batched step), the others are done
    k_batch_halted<<<(nsys + kBlock - 1) / kBlock, kBlock, 0, st>>>(b->d_ctls, nsys, b->d_halted);
    ATVH_HIP(hipGetLastError());
    std::vector<int> halted_h(nsys);
    ATVH_HIP(hipStreamSynchronize(st));
    ATVH_HIP(hipMemcpy(halted_h.data(), b->d_halted, nsys * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < nsys; ++k)
        if (halted_h[k])
            md_recover_and_finish(b->sys[k], nsteps, precision, energy_every);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_batch_exchange(atvh_batch* b, int64_t round, uint64_t seed, int* accepted) {
    if (!b)
        return ATVH_E_ARG;
    atvh_ctx* c = b->ctx;
    API_TRY
    if (round < 0)
        throw Error(ATVH_E_ARG, "atvh_batch_exchange: round < 0");
    for (atvh_system* s : b->sys) {
        if (s->n != b->sys[0]->n)
            throw Error(ATVH_E_ARG, "atvh_batch_exchange: replicas must hold the same atom count (an exchange "
                                    "swaps one configuration for another)");
        if (!s->integ_configured || s->thermostat != 1)
            throw Error(ATVH_E_STATE, "atvh_batch_exchange: every replica needs the Nose-Hoover thermostat");
    }
    ATVH_HIP(hipSetDevice(c->device));
    const int nsys = static_cast<int>(b->sys.size());
    k_batch_exchange<<<1, 64, 0, c->stream>>>(b->d_ctls, b->d_slot_member, nsys, static_cast<int>(round & 1),
                                              static_cast<unsigned long long>(seed),
                                              static_cast<unsigned long long>(round), b->d_accept);
    ATVH_HIP(hipGetLastError());
    if (accepted) {
        ATVH_HIP(hipMemcpyAsync(accepted, b->d_accept, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        ATVH_HIP(hipStreamSynchronize(c->stream));
    }
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_batch_ladder(atvh_batch* b, int* member_of_slot, double* T_of_member) {
    if (!b)
        return ATVH_E_ARG;
    atvh_ctx* c = b->ctx;
    API_TRY
    const int nsys = static_cast<int>(b->sys.size());
    ATVH_HIP(hipStreamSynchronize(c->stream));
    if (member_of_slot)
        ATVH_HIP(hipMemcpy(member_of_slot, b->d_slot_member, nsys * sizeof(int), hipMemcpyDeviceToHost));
    if (T_of_member)
        for (int k = 0; k < nsys; ++k) {
            DevCtl h{};
            ATVH_HIP(hipMemcpy(&h, b->sys[k]->ctl, sizeof(h), hipMemcpyDeviceToHost));
            T_of_member[k] = h.T_target;
        }
    return ATVH_OK;
    API_CATCH(c)
}

} // extern "C"
