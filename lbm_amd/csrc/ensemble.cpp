// Ensembles (lbm_ensemble_* in include/lbm_hip.h): many single-domain contexts advanced together
// by one launch of the batched step kernel per step, with the step sequence captured in HIP
// graphs and the launch shape chosen by a measuring autotuner.  The reference runs one domain
// per process; each member here behaves exactly as lbm_step on its own context would (same
// kernels, same arithmetic), which is what the tests check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "ctx.h"
#include "impl.h"

using namespace lbm;
using namespace lbm_host;

namespace {
constexpr int kGraphSteps = 16;    // steps per captured graph
constexpr int kChunk = 1024;       // steps per upload of the per-step inflow values
}  // namespace

struct lbm_ens {
    uint32_t tag = kEnsTag;
    std::vector<lbm_ctx*> mem;
    const Ops* ops = nullptr;
    int bgk3 = 0, hrr = 0, vre = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr;
    std::vector<hipEvent_t> ev_mem;
    int cus = 256;
    // launch shape
    int block = 256;
    int bpc = 0;                   // persistent grid of bpc blocks per CU; 0: one block per tile
    // device tables and their host images
    DevBuf tab, start, vals, clock, backup;
    std::vector<unsigned char> h_tab, h_vals;
    std::vector<int> h_start;
    int h_clock[3] = {0, 0, 0};
    EnsLaunch L{};
    // captured graphs by number of steps, valid for the launch in `key`
    std::map<int, hipGraphExec_t> graphs;
    std::vector<long long> key;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    size_t n_ev = 0;
    int timed_steps = 0;
    std::string err;
};

namespace {

// LBM_OK for a live ensemble handle; otherwise LBM_EINVAL with the message kept for
// lbm_ensemble_last_error(NULL) (and lbm_last_error(NULL))
int check_ens(const lbm_ens* e) {
    if (!e) return fail(nullptr, LBM_EINVAL, "null ensemble");
    if (e->tag != kEnsTag)
        return fail(nullptr, LBM_EINVAL, "not a live lbm_ens handle (a context handle, or a destroyed ensemble)");
    return LBM_OK;
}

int efail(lbm_ens* e, int code, const std::string& msg) {
    e->err = msg;
    return code;
}

#define EHIP(e, call)                                                                        \
    do {                                                                                     \
        hipError_t r_ = (call);                                                              \
        if (r_ != hipSuccess) return efail((e), LBM_EHIP, std::string(#call) + ": " + hipGetErrorString(r_)); \
    } while (0)

int ealloc(lbm_ens* e, DevBuf& b, size_t n, bool* moved) {
    if (b.n >= n && b.p) return LBM_OK;
    if (b.p) {
        EHIP(e, hipStreamSynchronize(e->stream));
        EHIP(e, hipFree(b.p));
        b.p = nullptr;
        b.n = 0;
    }
    EHIP(e, hipMalloc(&b.p, n));
    b.n = n;
    if (moved) *moved = true;
    return LBM_OK;
}

void drop_graphs(lbm_ens* e) {
    for (auto& g : e->graphs) hipGraphExecDestroy(g.second);
    e->graphs.clear();
}

int closed_axes(const lbm_desc& d) {
    return (d.ibnd > 10) + (d.jbnd > 10) + (d.kbnd > 10);
}

// Member tables for the current launch shape; drops the graphs when a kernel argument changes.
int build_tables(lbm_ens* e) {
    const int n = (int)e->mem.size();
    const size_t mb = e->ops->member_bytes;
    e->h_tab.assign(mb * n, 0);
    e->h_start.assign(n, 0);
    long long total = 0;
    for (int m = 0; m < n; ++m) {
        lbm_ctx* c = e->mem[m];
        if (c->avg_on) return efail(e, LBM_ESTATE, "member " + std::to_string(m) + ": time averaging is not batched; switch it off");
        MemberHost mh{};
        mh.lay = c->lay;
        mh.h = step_host(c, true);
        mh.buf[0] = c->buf[0];
        mh.buf[1] = c->buf[1];
        mh.cur0 = c->cur;
        mh.nrturb = c->nrturb;
        for (int q = 0; q < 3; ++q) mh.turb[q] = c->has_turb ? c->turb[q].p : nullptr;
        mh.shear = c->shear.p;
        mh.rho0 = c->d.rho0;
        mh.inflow = c->d.ibnd == 1;
        int tiles = 0;
        e->ops->ens_pack(mh, e->block, e->h_tab.data() + mb * m, &tiles);
        e->h_start[m] = (int)total;
        total += tiles;
    }
    if (total >= (1LL << 31)) return efail(e, LBM_EINVAL, "ensemble too large for 32-bit tile indices");
    bool moved = false;
    if (int rc = ealloc(e, e->tab, mb * n, &moved)) return rc;
    if (int rc = ealloc(e, e->start, sizeof(int) * n, &moved)) return rc;
    if (int rc = ealloc(e, e->clock, sizeof(e->h_clock), &moved)) return rc;
    if (int rc = ealloc(e, e->vals, (size_t)kChunk * n * 3 * e->ops->real_bytes, &moved)) return rc;
    EHIP(e, hipMemcpyAsync(e->tab.p, e->h_tab.data(), mb * n, hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipMemcpyAsync(e->start.p, e->h_start.data(), sizeof(int) * n, hipMemcpyHostToDevice, e->stream));
    EnsLaunch& L = e->L;
    L.tab = e->tab.p;
    L.start = static_cast<const int*>(e->start.p);
    L.nmem = n;
    L.total = (int)total;
    L.block = e->block;
    L.grid = e->bpc > 0 ? (int)std::min<long long>(total, (long long)e->bpc * e->cus) : (int)total;
    L.vals = e->vals.p;
    L.clock = static_cast<int*>(e->clock.p);
    L.bgk3 = e->bgk3;
    L.hrr = e->hrr;
    L.vre = e->vre;
    std::vector<long long> key = {(long long)(size_t)L.tab, (long long)(size_t)L.start, L.nmem, L.total, L.block,
                                  L.grid, (long long)(size_t)L.vals, (long long)(size_t)L.clock};
    if (moved || key != e->key) {
        drop_graphs(e);
        e->key = key;
    }
    return LBM_OK;
}

int graph_for(lbm_ens* e, int steps, hipGraphExec_t* out) {
    auto it = e->graphs.find(steps);
    if (it != e->graphs.end()) {
        *out = it->second;
        return LBM_OK;
    }
    EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int s = 0; s < steps && le == hipSuccess; ++s) le = e->ops->ens_step(e->L, e->stream);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (le != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return efail(e, LBM_EHIP, std::string("batched step launch: ") + hipGetErrorString(le));
    }
    if (ce != hipSuccess) return efail(e, LBM_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    hipGraphExec_t x = nullptr;
    const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ie != hipSuccess) return efail(e, LBM_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    e->graphs[steps] = x;
    *out = x;
    return LBM_OK;
}

hipEvent_t next_ev(lbm_ens* e) {
    if (e->n_ev == e->ev.size()) {
        hipEvent_t x;
        if (hipEventCreate(&x) != hipSuccess) return nullptr;
        e->ev.push_back(x);
    }
    return e->ev[e->n_ev++];
}

// Batched steps it0+1 .. it0+nsteps of members that hold a collided state.
int batched(lbm_ens* e, int it0, int nsteps) {
    const int n = (int)e->mem.size();
    const int rb = e->ops->real_bytes;
    EHIP(e, hipStreamSynchronize(e->stream));    // host images may be rewritten
    if (int rc = build_tables(e)) return rc;
    int done = 0;
    while (done < nsteps) {
        const int chunk = std::min(kChunk, nsteps - done);
        if (done > 0) EHIP(e, hipStreamSynchronize(e->stream));
        e->h_vals.assign((size_t)chunk * n * 3 * rb, 0);
        for (int s = 0; s < chunk; ++s)
            for (int m = 0; m < n; ++m) {
                if (e->mem[m]->d.ibnd != 1) continue;
                double v[3];
                inflow_values(e->mem[m], it0 + done + s + 1, &v[0], &v[1], &v[2]);
                for (int q = 0; q < 3; ++q) {
                    const size_t o = ((size_t)s * n + m) * 3 + q;
                    if (rb == 8) reinterpret_cast<double*>(e->h_vals.data())[o] = v[q];
                    else reinterpret_cast<float*>(e->h_vals.data())[o] = (float)v[q];
                }
            }
        EHIP(e, hipMemcpyAsync(e->vals.p, e->h_vals.data(), e->h_vals.size(), hipMemcpyHostToDevice, e->stream));
        e->h_clock[0] = it0 + done + 1;
        e->h_clock[1] = done;   // buffer parity: steps since the members' current buffers were taken
        e->h_clock[2] = done;   // the uploaded table's first step
        EHIP(e, hipMemcpyAsync(e->clock.p, e->h_clock, sizeof(e->h_clock), hipMemcpyHostToDevice, e->stream));
        hipEvent_t t0 = nullptr;
        if (e->timing && (t0 = next_ev(e))) hipEventRecord(t0, e->stream);
        int left = chunk;
        while (left > 0) {
            const int g = std::min(left, kGraphSteps);
            hipGraphExec_t x;
            if (int rc = graph_for(e, g, &x)) return rc;
            EHIP(e, hipGraphLaunch(x, e->stream));
            left -= g;
        }
        if (e->timing && t0) {
            hipEvent_t t1 = next_ev(e);
            if (t1) hipEventRecord(t1, e->stream);
            e->timed_steps += chunk;
        }
        done += chunk;
    }
    for (lbm_ctx* c : e->mem)
        for (int s = 0; s < nsteps; ++s) {
            c->cur ^= 1;
            c->ghost_raw[c->cur] = false;
        }
    return LBM_OK;
}

}  // namespace

extern "C" {

const char* lbm_ensemble_last_error(const lbm_ens* e) {
    return e && e->tag == kEnsTag ? e->err.c_str() : lbm_last_error(nullptr);
}

int lbm_ensemble_create(lbm_ctx* const* members, int n, lbm_ens** out) {
    if (!out) return fail(nullptr, LBM_EINVAL, "null argument");
    *out = nullptr;
    if (!members || n < 1) return fail(nullptr, LBM_EINVAL, "no members");
    for (int m = 0; m < n; ++m)
        if (check_ctx(members[m]) != LBM_OK)
            return fail(nullptr, LBM_EINVAL, "member " + std::to_string(m) + " is not a live lbm_ctx handle");
    lbm_ens* e = new lbm_ens();
    auto bad = [&](int code, const std::string& msg) {
        // keep the message reachable: the ensemble is returned so lbm_ensemble_last_error works
        e->err = msg;
        *out = e;
        return code;
    };
    for (int m = 0; m < n; ++m) {
        lbm_ctx* c = members[m];
        e->mem.push_back(c);
        const lbm_desc& d = c->d;
        const lbm_desc& d0 = members[0]->d;
        if (c->ops != members[0]->ops || d.q != d0.q || d.ibgk != d0.ibgk || d.ihrr != d0.ihrr || d.ivreman != d0.ivreman)
            return bad(LBM_EINVAL, "member " + std::to_string(m) + ": all members need the same lattice, storage and collision model");
        if (c->ops->raw_bytes > 8)
            return bad(LBM_EINVAL, "member " + std::to_string(m) + ": ensembles take FP32, FP64, FP16 or BF16 storage");
        if (d.jbnd == LBM_BND_HALO)
            return bad(LBM_EINVAL, "member " + std::to_string(m) + ": halo domains are stepped on their own");
        if (c->turb_on || c->buoy_on)
            return bad(LBM_EINVAL, "member " + std::to_string(m) + ": members with turbines or buoyancy are stepped on their own");
        if (closed_axes(d) > 1 || (closed_axes(d) == 1 && c->fold_dir < 0))
            return bad(LBM_EINVAL, "member " + std::to_string(m) +
                                       ": ensembles take members with at most one closed axis (its face pass runs inside the step kernel)");
        for (int q = 0; q < m; ++q)
            if (members[q] == c) return bad(LBM_EINVAL, "member " + std::to_string(m) + " repeats member " + std::to_string(q));
    }
    const lbm_desc& d0 = members[0]->d;
    e->ops = members[0]->ops;
    e->bgk3 = d0.ibgk == 3;
    e->hrr = d0.ihrr == 1;
    e->vre = d0.ivreman == 1;
    hipError_t r;
    int dev = 0;
    if ((r = hipGetDevice(&dev)) != hipSuccess || (r = hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
        (r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess ||
        (r = hipEventCreateWithFlags(&e->ev_done, hipEventDisableTiming)) != hipSuccess)
        return bad(LBM_EHIP, std::string("ensemble setup: ") + hipGetErrorString(r));
    e->ev_mem.resize(n, nullptr);
    for (int m = 0; m < n; ++m)
        if ((r = hipEventCreateWithFlags(&e->ev_mem[m], hipEventDisableTiming)) != hipSuccess)
            return bad(LBM_EHIP, std::string("ensemble setup: ") + hipGetErrorString(r));
    *out = e;
    return LBM_OK;
}

int lbm_ensemble_destroy(lbm_ens* e) {
    if (!e) return LBM_OK;
    if (int rc = check_ens(e)) return rc;
    e->tag = 0;
    if (e->stream) hipStreamSynchronize(e->stream);
    drop_graphs(e);
    for (DevBuf* b : {&e->tab, &e->start, &e->vals, &e->clock, &e->backup}) dev_free(*b);
    for (hipEvent_t x : e->ev) hipEventDestroy(x);
    for (hipEvent_t x : e->ev_mem)
        if (x) hipEventDestroy(x);
    if (e->ev_done) hipEventDestroy(e->ev_done);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return LBM_OK;
}

int lbm_ensemble_set_launch(lbm_ens* e, int block, int blocks_per_cu) {
    if (int rc_ = check_ens(e)) return rc_;
    if (block < 64 || block > 1024 || block % 64 != 0) return efail(e, LBM_EINVAL, "block must be a multiple of 64 in [64, 1024]");
    if (blocks_per_cu < 0 || blocks_per_cu > 64) return efail(e, LBM_EINVAL, "blocks_per_cu must be in [0, 64]");
    e->block = block;
    e->bpc = blocks_per_cu;
    return LBM_OK;
}

int lbm_ensemble_get_launch(lbm_ens* e, int* block, int* blocks_per_cu) {
    if (int rc_ = check_ens(e)) return rc_;
    if (!block || !blocks_per_cu) return efail(e, LBM_EINVAL, "null argument");
    *block = e->block;
    *blocks_per_cu = e->bpc;
    return LBM_OK;
}

int lbm_ensemble_step(lbm_ens* e, int it0, int nsteps) {
    if (int rc_ = check_ens(e)) return rc_;
    if (nsteps < 0) return efail(e, LBM_EINVAL, "nsteps < 0");
    if (nsteps == 0) return LBM_OK;
    const int n = (int)e->mem.size();
    bool first = false;
    for (int m = 0; m < n; ++m) {
        lbm_ctx* c = e->mem[m];
        if (!c->have_state) return efail(e, LBM_ESTATE, "member " + std::to_string(m) + " has no state");
        first = first || c->pending || c->use_ext;
    }
    if (first) {
        // the first step after set_state / init_equilibrium collides in place: run it per member
        for (int m = 0; m < n; ++m)
            if (int rc = lbm_step(e->mem[m], it0, 1))
                return efail(e, rc, "member " + std::to_string(m) + ": " + e->mem[m]->err);
        ++it0;
        if (--nsteps == 0) return LBM_OK;
    }
    for (int m = 0; m < n; ++m) {
        EHIP(e, hipEventRecord(e->ev_mem[m], e->mem[m]->stream));
        EHIP(e, hipStreamWaitEvent(e->stream, e->ev_mem[m], 0));
    }
    if (int rc = batched(e, it0, nsteps)) return rc;
    EHIP(e, hipEventRecord(e->ev_done, e->stream));
    for (int m = 0; m < n; ++m) EHIP(e, hipStreamWaitEvent(e->mem[m]->stream, e->ev_done, 0));
    return LBM_OK;
}

int lbm_ensemble_synchronize(lbm_ens* e) {
    if (int rc_ = check_ens(e)) return rc_;
    EHIP(e, hipStreamSynchronize(e->stream));
    return LBM_OK;
}

int lbm_ensemble_set_timing(lbm_ens* e, int on) {
    if (int rc_ = check_ens(e)) return rc_;
    e->timing = on != 0;
    return LBM_OK;
}

int lbm_ensemble_timing_read(lbm_ens* e, double* ms, int* nsteps) {
    if (int rc_ = check_ens(e)) return rc_;
    if (!ms || !nsteps) return efail(e, LBM_EINVAL, "null argument");
    EHIP(e, hipStreamSynchronize(e->stream));
    double t = 0.0;
    for (size_t q = 0; q + 1 < e->n_ev; q += 2) {
        float x = 0.f;
        EHIP(e, hipEventElapsedTime(&x, e->ev[q], e->ev[q + 1]));
        t += x;
    }
    *ms = t;
    *nsteps = e->timed_steps;
    e->n_ev = 0;
    e->timed_steps = 0;
    return LBM_OK;
}

// Autotuner: every candidate (block, blocks per CU) advances the ensemble by `warm` untimed and
// `trial_steps` timed steps from the same saved state; the fastest is kept and the state and the
// members' bookkeeping are restored afterwards.
int lbm_ensemble_tune(lbm_ens* e, int it0, int trial_steps, int* best_block, int* best_bpc, double* best_ms) {
    if (int rc_ = check_ens(e)) return rc_;
    if (trial_steps < 1) return efail(e, LBM_EINVAL, "trial_steps must be >= 1");
    const int n = (int)e->mem.size();
    struct Saved {
        int cur;
        bool pending, use_ext, raw0, raw1;
    };
    std::vector<Saved> sv(n);
    std::vector<size_t> off(n + 1, 0);
    for (int m = 0; m < n; ++m) {
        lbm_ctx* c = e->mem[m];
        if (!c->have_state) return efail(e, LBM_ESTATE, "member " + std::to_string(m) + " has no state");
        sv[m] = {c->cur, c->pending, c->use_ext, c->ghost_raw[0], c->ghost_raw[1]};
        off[m + 1] = off[m] + 2 * c->buf_bytes + c->tau.n;
    }
    for (int m = 0; m < n; ++m) {
        EHIP(e, hipEventRecord(e->ev_mem[m], e->mem[m]->stream));
        EHIP(e, hipStreamWaitEvent(e->stream, e->ev_mem[m], 0));
    }
    if (int rc = ealloc(e, e->backup, off[n], nullptr)) return rc;
    char* bk = static_cast<char*>(e->backup.p);
    auto copy_all = [&](bool save) -> int {
        for (int m = 0; m < n; ++m) {
            lbm_ctx* c = e->mem[m];
            char* p = bk + off[m];
            void* parts[3] = {c->buf[0], c->buf[1], c->tau.p};
            const size_t sizes[3] = {c->buf_bytes, c->buf_bytes, c->tau.n};
            for (int q = 0; q < 3; ++q) {
                if (!sizes[q]) continue;
                if (save) EHIP(e, hipMemcpyAsync(p, parts[q], sizes[q], hipMemcpyDeviceToDevice, e->stream));
                else EHIP(e, hipMemcpyAsync(parts[q], p, sizes[q], hipMemcpyDeviceToDevice, e->stream));
                p += sizes[q];
            }
            if (!save) {
                c->cur = sv[m].cur;
                c->pending = sv[m].pending;
                c->use_ext = sv[m].use_ext;
                c->ghost_raw[0] = sv[m].raw0;
                c->ghost_raw[1] = sv[m].raw1;
            }
        }
        EHIP(e, hipStreamSynchronize(e->stream));
        return LBM_OK;
    };
    if (int rc = copy_all(true)) return rc;
    const bool timing = e->timing;
    const int blocks[] = {64, 128, 256, 512, 1024};
    const int bpcs[] = {0, 1, 2, 4, 8};
    double best = -1.0;
    int bb = e->block, bp = e->bpc;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    EHIP(e, hipEventCreate(&t0));
    EHIP(e, hipEventCreate(&t1));
    int rc = LBM_OK;
    for (int b : blocks) {
        for (int p : bpcs) {
            e->block = b;
            e->bpc = p;
            e->timing = false;
            if ((rc = lbm_ensemble_step(e, it0, 2))) break;            // first step, graph capture
            if (hipEventRecord(t0, e->stream) != hipSuccess) { rc = efail(e, LBM_EHIP, "hipEventRecord"); break; }
            if ((rc = lbm_ensemble_step(e, it0 + 2, trial_steps))) break;
            if (hipEventRecord(t1, e->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) {
                rc = efail(e, LBM_EHIP, "timing event");
                break;
            }
            float ms = 0.f;
            hipEventElapsedTime(&ms, t0, t1);
            if (best < 0 || ms < best) {
                best = ms;
                bb = b;
                bp = p;
            }
            if ((rc = copy_all(false))) break;
        }
        if (rc) break;
    }
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    e->timing = timing;
    e->block = bb;
    e->bpc = bp;
    if (rc) {
        copy_all(false);
        return rc;
    }
    if ((rc = copy_all(false))) return rc;
    EHIP(e, hipEventRecord(e->ev_done, e->stream));
    for (int m = 0; m < n; ++m) EHIP(e, hipStreamWaitEvent(e->mem[m]->stream, e->ev_done, 0));
    if (best_block) *best_block = bb;
    if (best_bpc) *best_bpc = bp;
    if (best_ms) *best_ms = best / trial_steps;
    return LBM_OK;
}

}  // extern "C"
