// ctx_system.hip — context, device frame (upload/download), timing, C-ABI glue.
//
// Replaces cuda::Backend::init / materialize / retrieve
// (src/cuda/driver/driver.cu:25-86,177) and cpu::to_cuda / cuda::to_cpu
// (src/cuda/frame/transfer.cu:145-475).

#include "atvh_internal.hpp"

#include <cmath>
#include <cstring>

using namespace atvh;

namespace atvh {

hipEvent_t take_event(atvh_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    ATVH_HIP(hipEventCreate(&e));
    return e;
}

bool timing_on(const atvh_ctx* c, const char* nm) {
    return c->timing && (c->timing_only.empty() || c->timing_only == nm);
}

ScopedTimer::ScopedTimer(atvh_ctx* c, const char* nm) : ctx(c), name(nm) {
    if (timing_on(ctx, nm)) {
        start = take_event(ctx);
        ATVH_HIP(hipEventRecord(start, ctx->stream));
    }
}

ScopedTimer::~ScopedTimer() {
    if (!start)
        return;
    hipEvent_t stop = take_event(ctx);
    if (hipEventRecord(stop, ctx->stream) == hipSuccess)
        ctx->timers[name].pending.emplace_back(start, stop);
}

static void resolve_timers(atvh_ctx* c) {
    ATVH_HIP(hipStreamSynchronize(c->stream));
    for (auto& kv : c->timers) {
        for (auto& pr : kv.second.pending) {
            float ms = 0.f;
            ATVH_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            kv.second.total_ms += ms;
            kv.second.launches += 1;
            c->event_pool.push_back(pr.first);
            c->event_pool.push_back(pr.second);
        }
        kv.second.pending.clear();
    }
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

namespace {

// Frame-order gather for downloads; positions are wrapped into [0, L) here
// (wrap_orthorhombic, pbc.h:37-51), velocities get the pending NH scale.
__global__ void k_gather_out(int64_t n, const int* __restrict__ tag, const double* __restrict__ a,
                             const double* __restrict__ b, const double* __restrict__ c,
                             double* __restrict__ oa, double* __restrict__ ob,
                             double* __restrict__ oc, const DevCtl* ctl, int mode, double la,
                             double lb, double lc) {
    const double s = mode == 1 ? ctl->nh_scale : 1.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int t = tag[i];
        double va = a[i], vb = b[i], vc = c[i];
        if (mode == 1) {
            va *= s;
            vb *= s;
            vc *= s;
        } else if (mode == 2) {
            va -= la * floor(va / la);
            vb -= lb * floor(vb / lb);
            vc -= lc * floor(vc / lc);
        }
        oa[t] = va;
        ob[t] = vb;
        oc[t] = vc;
    }
}

__global__ void k_iota(int64_t n, int* __restrict__ tag) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        tag[i] = static_cast<int>(i);
}

void sys_free_all(atvh_system* s) {
    double** dptrs[] = {&s->x,  &s->y,  &s->z,  &s->vx,  &s->vy,  &s->vz,  &s->fx, &s->fy,
                        &s->fz, &s->q,  &s->ax, &s->ay,  &s->az,  &s->avx, &s->avy, &s->avz,
                        &s->afx, &s->afy, &s->afz, &s->aq, &s->xs, &s->ys, &s->zs, &s->partials,
                        &s->d_thermo};
    for (auto p : dptrs)
        dev_free(*p);
    dev_free(s->type);
    dev_free(s->tag);
    dev_free(s->atype);
    dev_free(s->atag);
    dev_free(s->ctl);
    nl_free(s);
    pair_free(s);
    coul_free(s);
    npt_free(s);
}

} // namespace

extern "C" {

int atvh_ctx_create(int device, atvh_ctx** out) {
    if (!out)
        return ATVH_E_ARG;
    *out = nullptr;
    auto* c = new atvh_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return ATVH_E_HIP;
    }
    c->stream = c->own_stream;
    try {
        noalloc_bind(device);
    } catch (const Error&) {
        (void)hipStreamDestroy(c->own_stream);
        delete c;
        return ATVH_E_HIP;
    }
    *out = c;
    return ATVH_OK;
}

int atvh_ctx_destroy(atvh_ctx* c) {
    if (!c)
        return ATVH_OK;
    if (c->live_systems > 0) {
        c->err = "atvh_ctx_destroy: systems created on this context are still alive";
        return ATVH_E_STATE;
    }
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->timers)
        for (auto& pr : kv.second.pending) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (auto e : c->event_pool)
        (void)hipEventDestroy(e);
    if (c->own_stream)
        (void)hipStreamDestroy(c->own_stream);
    delete c;
    return ATVH_OK;
}

const char* atvh_last_error(const atvh_ctx* c) { return c ? c->err.c_str() : "null context"; }

int atvh_ctx_set_stream(atvh_ctx* c, void* stream) {
    if (!c)
        return ATVH_E_ARG;
    c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
    return ATVH_OK;
}

void* atvh_ctx_stream(atvh_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int atvh_ctx_synchronize(atvh_ctx* c) {
    if (!c)
        return ATVH_E_ARG;
    API_TRY
    ATVH_HIP(hipStreamSynchronize(c->stream));
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_ctx_enable_timing(atvh_ctx* c, int enable) {
    if (!c)
        return ATVH_E_ARG;
    c->timing = enable != 0;
    return ATVH_OK;
}

int atvh_ctx_timing_only(atvh_ctx* c, const char* name) {
    if (!c)
        return ATVH_E_ARG;
    API_TRY
    c->timing_only = name ? name : "";
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_ctx_kernel_time(atvh_ctx* c, const char* name, double* total_ms, int64_t* launches) {
    if (!c || !name)
        return ATVH_E_ARG;
    API_TRY
    resolve_timers(c);
    auto it = c->timers.find(name);
    if (total_ms)
        *total_ms = it == c->timers.end() ? 0.0 : it->second.total_ms;
    if (launches)
        *launches = it == c->timers.end() ? 0 : it->second.launches;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_ctx_reset_timing(atvh_ctx* c) {
    if (!c)
        return ATVH_E_ARG;
    API_TRY
    resolve_timers(c);
    for (auto& kv : c->timers) {
        kv.second.total_ms = 0;
        kv.second.launches = 0;
    }
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_ctx_mem_info(atvh_ctx* c, int64_t* free_bytes, int64_t* total_bytes) {
    if (!c)
        return ATVH_E_ARG;
    API_TRY
    ATVH_HIP(hipSetDevice(c->device));
    size_t f = 0, t = 0;
    ATVH_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes)
        *free_bytes = static_cast<int64_t>(f);
    if (total_bytes)
        *total_bytes = static_cast<int64_t>(t);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_system_create(atvh_ctx* c, int64_t n, const double box[3], atvh_system** out) {
    if (!c || !out || !box || n <= 0)
        return ATVH_E_ARG;
    *out = nullptr;
    auto* s = new atvh_system();
    API_TRY
    if (n > (int64_t(1) << 30))
        throw Error(ATVH_E_ARG, "atvh_system_create: n exceeds 2^30 atoms");
    for (int d = 0; d < 3; ++d)
        if (!(box[d] > 0.0) || !std::isfinite(box[d]))
            throw Error(ATVH_E_ARG, "atvh_system_create: box lengths must be positive");
    ATVH_HIP(hipSetDevice(c->device));
    s->ctx = c;
    s->n = n;
    s->npad = (n + 63) / 64 * 64;
    std::memcpy(s->box, box, sizeof(double) * 3);
    double** arrs[] = {&s->x,  &s->y,  &s->z,   &s->vx,  &s->vy,  &s->vz,  &s->fx, &s->fy, &s->fz,
                       &s->q,  &s->ax, &s->ay,  &s->az,  &s->avx, &s->avy, &s->avz, &s->afx,
                       &s->afy, &s->afz, &s->aq, &s->xs, &s->ys, &s->zs};
    for (auto p : arrs)
        dev_alloc(p, n);
    dev_alloc(&s->type, n);
    dev_alloc(&s->tag, n);
    dev_alloc(&s->atype, n);
    dev_alloc(&s->atag, n);
    dev_alloc(&s->ctl, 1);
    s->npartials = kMaxGrid;
    dev_alloc(&s->partials, size_t(2 * s->npartials) * kPartialStride); // pair | kinetic
    DevCtl h{};
    h.nh_scale = 1.0;
    h.always = 1;
    ATVH_HIP(hipMemcpyAsync(s->ctl, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    for (auto p : arrs)
        ATVH_HIP(hipMemsetAsync(*p, 0, n * sizeof(double), c->stream));
    ATVH_HIP(hipMemsetAsync(s->type, 0, n * sizeof(int), c->stream));
    k_iota<<<grid_capped(n), kBlock, 0, c->stream>>>(n, s->tag);
    ATVH_HIP(hipGetLastError());
    ATVH_HIP(hipStreamSynchronize(c->stream));
    c->live_systems += 1;
    *out = s;
    return ATVH_OK;
    }
    catch (const atvh::Error& e_) {
        c->err = e_.what();
        sys_free_all(s);
        delete s;
        return e_.code;
    }
    catch (const std::exception& e_) {
        c->err = e_.what();
        sys_free_all(s);
        delete s;
        return ATVH_E_ARG;
    }
}

int atvh_system_destroy(atvh_system* s) {
    if (!s)
        return ATVH_OK;
    if (s->ctx) {
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
    }
    sys_free_all(s);
    if (s->ctx)
        s->ctx->live_systems -= 1;
    delete s;
    return ATVH_OK;
}

int64_t atvh_system_natoms(const atvh_system* s) { return s ? s->n : 0; }

int atvh_system_memory(const atvh_system* s, int part, int64_t* bytes) {
    if (!s || !bytes)
        return ATVH_E_ARG;
    const int64_t n = s->n;
    switch (part) {
    case ATVH_MEM_FRAME:
        // 23 FP64 per-atom columns (state, staging, x_save), 4 int32 columns,
        // the reduction partials and the control block
        *bytes = n * (23 * 8 + 4 * 4) + int64_t(2) * s->npartials * kPartialStride * 8 + int64_t(sizeof(DevCtl));
        return ATVH_OK;
    case ATVH_MEM_NLIST:
        *bytes = nl_device_bytes(s);
        return ATVH_OK;
    case ATVH_MEM_PAIR:
        *bytes = pair_device_bytes(s);
        return ATVH_OK;
    case ATVH_MEM_PME:
        *bytes = pme_device_bytes(s);
        return ATVH_OK;
    case ATVH_MEM_INTEGRATOR:
        *bytes = int64_t(2) * s->thermo_cap * 8; // the per-step energy record
        return ATVH_OK;
    default:
        return ATVH_E_ARG;
    }
}

int atvh_system_box(const atvh_system* s, double box[3]) {
    if (!s || !box)
        return ATVH_E_ARG;
    std::memcpy(box, s->box, sizeof(double) * 3);
    return ATVH_OK;
}

int atvh_frame_upload(atvh_system* s, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const double* fx,
                      const double* fy, const double* fz, const double* mass, const int* type,
                      const double* charge) {
    if (!s || !x || !y || !z || !mass)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    ATVH_HIP(hipSetDevice(c->device));
    const int64_t n = s->n;
    // per-type masses (the integrators read mass by type)
    int ntypes = 1;
    double mt[kMaxTypes];
    bool seen[kMaxTypes] = {false};
    for (int64_t i = 0; i < n; ++i) {
        const int t = type ? type[i] : 0;
        if (t < 0 || t >= kMaxTypes)
            throw Error(ATVH_E_ARG, "atvh_frame_upload: type out of range [0,16)");
        if (!(mass[i] > 0.0))
            throw Error(ATVH_E_ARG, "atvh_frame_upload: mass must be positive");
        if (!seen[t]) {
            seen[t] = true;
            mt[t] = mass[i];
        } else if (mt[t] != mass[i]) {
            throw Error(ATVH_E_ARG, "atvh_frame_upload: atoms of one type must share one mass");
        }
        if (t + 1 > ntypes)
            ntypes = t + 1;
    }
    for (int t = 0; t < ntypes; ++t)
        s->mass_t[t] = seen[t] ? mt[t] : 1.0;
    s->ntypes = ntypes;
    hipStream_t st = c->stream;
    auto up = [&](double* d, const double* h) {
        if (h)
            ATVH_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, st));
        else
            ATVH_HIP(hipMemsetAsync(d, 0, n * sizeof(double), st));
    };
    up(s->x, x);
    up(s->y, y);
    up(s->z, z);
    up(s->vx, vx);
    up(s->vy, vy);
    up(s->vz, vz);
    up(s->fx, fx);
    up(s->fy, fy);
    up(s->fz, fz);
    up(s->q, charge);
    s->has_charge = charge != nullptr;
    // charge sums in ascending atom order (KSpacePME self / net-charge terms,
    // pme.cpp:222-226) and the bound of any spread grid point
    s->q_sum = s->q_sq = s->q_abs = s->q_max = 0.0;
    if (charge)
        for (int64_t i = 0; i < n; ++i) {
            s->q_sq += charge[i] * charge[i];
            s->q_sum += charge[i];
            s->q_abs += std::fabs(charge[i]);
            s->q_max = std::max(s->q_max, std::fabs(charge[i]));
        }
    if (type)
        ATVH_HIP(hipMemcpyAsync(s->type, type, n * sizeof(int), hipMemcpyHostToDevice, st));
    else
        ATVH_HIP(hipMemsetAsync(s->type, 0, n * sizeof(int), st));
    k_iota<<<grid_capped(n), kBlock, 0, st>>>(n, s->tag);
    ATVH_HIP(hipGetLastError());
    // a fresh frame: no pending thermostat scale; list must be rebuilt
    ATVH_HIP(hipStreamSynchronize(st));
    DevCtl h{};
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    h.nh_scale = 1.0;
    h.always = 1;
    ATVH_HIP(hipMemcpyAsync(s->ctl, &h, sizeof(h), hipMemcpyHostToDevice, st));
    s->nl_built = false;
    ATVH_HIP(hipStreamSynchronize(st));
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_frame_download(atvh_system* s, double* x, double* y, double* z, double* vx, double* vy,
                        double* vz, double* fx, double* fy, double* fz) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    ATVH_HIP(hipSetDevice(c->device));
    const int64_t n = s->n;
    hipStream_t st = c->stream;
    auto fetch3 = [&](const double* a, const double* b, const double* cc, double* ha, double* hb,
                      double* hc, int mode) {
        if (!ha && !hb && !hc)
            return;
        k_gather_out<<<grid_capped(n), kBlock, 0, st>>>(n, s->tag, a, b, cc, s->ax, s->ay, s->az,
                                                        s->ctl, mode, s->box[0], s->box[1], s->box[2]);
        ATVH_HIP(hipGetLastError());
        if (ha)
            ATVH_HIP(hipMemcpyAsync(ha, s->ax, n * sizeof(double), hipMemcpyDeviceToHost, st));
        if (hb)
            ATVH_HIP(hipMemcpyAsync(hb, s->ay, n * sizeof(double), hipMemcpyDeviceToHost, st));
        if (hc)
            ATVH_HIP(hipMemcpyAsync(hc, s->az, n * sizeof(double), hipMemcpyDeviceToHost, st));
        ATVH_HIP(hipStreamSynchronize(st));
    };
    fetch3(s->x, s->y, s->z, x, y, z, 2);
    fetch3(s->vx, s->vy, s->vz, vx, vy, vz, 1);
    fetch3(s->fx, s->fy, s->fz, fx, fy, fz, 0);
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_energy(atvh_system* s, double* pe, double* virial6, double* ke) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    DevCtl h{};
    ATVH_HIP(hipStreamSynchronize(c->stream));
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    if (pe)
        *pe = h.pe;
    if (virial6)
        for (int k = 0; k < 6; ++k)
            virial6[k] = h.vir[k];
    if (ke)
        *ke = h.ke;
    return ATVH_OK;
    API_CATCH(c)
}

int atvh_thermostat_state(atvh_system* s, double* xi_out, const double* xi_in) {
    if (!s)
        return ATVH_E_ARG;
    atvh_ctx* c = s->ctx;
    API_TRY
    ATVH_HIP(hipStreamSynchronize(c->stream));
    DevCtl h{};
    ATVH_HIP(hipMemcpy(&h, s->ctl, sizeof(h), hipMemcpyDeviceToHost));
    if (xi_out)
        *xi_out = h.xi;
    if (xi_in) {
        h.xi = *xi_in;
        ATVH_HIP(hipMemcpy(s->ctl, &h, sizeof(h), hipMemcpyHostToDevice));
    }
    return ATVH_OK;
    API_CATCH(c)
}

} // extern "C"
