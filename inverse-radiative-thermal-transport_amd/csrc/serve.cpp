// serve.cpp -- request loop over design variants with continuous admission.
//
// The reference runs one scene at a time inside ThermalRenderer
// (thermal_renderer.cpp:958-1133: forward simulation, adjoint, gradient); this
// server runs many designs concurrently on one GPU.  Worker threads (slots)
// take requests in arrival order, first fit against the HBM budget, so a
// request that arrives while others run is admitted as soon as a slot and
// its footprint are free.  Each request owns a scene, and with it a HIP
// stream and stream-ordered device buffers, so the kernels of different
// requests overlap on the device and allocation does not synchronise it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "irtt_internal.hpp"

namespace irtt {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Req {
    uint64_t ticket = 0;
    // owned copies of the inputs
    std::vector<float> xyz, kelvin0, dxdp;
    std::vector<uint32_t> tri, vertex, param;
    std::vector<irtt_material> mat;
    std::vector<uint8_t> fixed;
    std::vector<double> target, source;
    irtt_request r{};  // scalars and output pointers
    uint32_t nb = 0;
    uint64_t footprint = 0;   // HBM counted for this request while it runs
    int state = IRTT_REQ_QUEUED;
    Clock::time_point t_submit, t_start, t_end;
};

// at most this many requests of one group share a scene and a batched solve
constexpr size_t kMaxGroup = 32;

bool same_bytes(const void *a, const void *b, size_t n) { return n == 0 || !std::memcmp(a, b, n); }

template <typename T>
bool same_vec(const std::vector<T> &a, const std::vector<T> &b)
{
    return a.size() == b.size() && same_bytes(a.data(), b.data(), a.size() * sizeof(T));
}

// requests a and b may share one scene, transport build and batched solves
bool shareable(const Req &a, const Req &b)
{
    const irtt_request &x = a.r, &y = b.r;
    return x.group && x.group == y.group && x.nv == y.nv && x.nt == y.nt &&
           x.bin_mode == y.bin_mode && x.operator_kind == y.operator_kind &&
           same_bytes(&x.transport, &y.transport, sizeof x.transport) &&
           same_bytes(&x.solver, &y.solver, sizeof x.solver) && x.solver.steady &&
           x.solver.jacobian == IRTT_JAC_EXACT && same_vec(a.xyz, b.xyz) &&
           same_vec(a.tri, b.tri) && same_vec(a.mat, b.mat) && same_vec(a.kelvin0, b.kelvin0) &&
           same_vec(a.fixed, b.fixed) && a.target.empty() == b.target.empty();
}

// HBM needed by one request: fixed-point + FP32 operator (+ FP16 copy), or
// for the sparse layout the peak of its build (hash table + the two key /
// value arrays being sorted, at most one entry per photon or bin), solver
// vectors and transposed-product partials, geometry and LBVH
uint64_t footprint(const irtt_request &q, uint32_t nb)
{
    const uint64_t ld = (nb + 63) / 64 * 64;
    uint64_t b = 0;
    if (q.transport.layout == IRTT_LAYOUT_SPARSE) {
        const uint64_t keys = (uint64_t)q.transport.threads_per_tri * q.transport.rays_per_thread *
                                  q.nt + nb;
        uint64_t cap = 1024;
        while (cap < keys + keys / 4) cap <<= 1;
        b = 16 * cap + 32 * keys;
    } else {
        b = (uint64_t)nb * ld * 12 + (((uint64_t)nb * ld) >> irtt::kTouchShift);
    }
    if (q.operator_kind == IRTT_OP_DENSE_F16) b += (uint64_t)nb * ld * 2 + nb * 4;
    b += ld * 8 * 40 + ((nb + 511) / 512) * ld * 8;
    b += (uint64_t)q.nt * 256 + (uint64_t)q.nv * 64 + (64ull << 20);
    if (q.target) {  // the adjoint trace's absorber record, kept paths and record buffers
        const uint64_t photons = (uint64_t)q.transport.threads_per_tri *
                                 q.transport.rays_per_thread * q.nt;
        b += adjoint_record_bytes(photons, photons) + kept_path_bytes(photons);
    }
    return b;
}

}  // namespace

struct Server {
    irtt_serve_config cfg{};
    uint64_t budget = 0, in_use = 0;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::vector<std::thread> workers;

    // first queued request whose footprint fits, with the queued requests
    // that can share its scene (same group, identical inputs; the first
    // one's footprint plus per-design vectors); empty if none fits
    std::vector<Req *> admit_locked()
    {
        std::vector<Req *> g;
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            if ((*it)->footprint > budget - in_use) continue;
            Req *q = *it;
            uint64_t need = q->footprint;
            g.push_back(q);
            if (q->r.group) {
                const uint64_t extra = 64ull * 8 * q->nb + (64ull << 10);
                for (auto jt = std::next(it); jt != queue.end() && g.size() < kMaxGroup; ++jt)
                    if (shareable(*q, **jt) && need + extra <= budget - in_use) {
                        g.push_back(*jt);
                        need += extra;
                    }
            }
            for (Req *r : g) {
                queue.erase(std::find(queue.begin(), queue.end(), r));
                r->state = IRTT_REQ_RUNNING;
                r->t_start = Clock::now();
            }
            // the group's memory is counted on its first request
            g[0]->footprint = need;
            for (size_t k = 1; k < g.size(); k++) g[k]->footprint = 0;
            in_use += need;
            return g;
        }
        return g;
    }

    void worker()
    {
        for (;;) {
            std::vector<Req *> g;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_work.wait(lk, [&] { return stop || !(g = admit_locked()).empty(); });
                if (g.empty()) return;  // stop with nothing admissible left
            }
            // a request is done once its results are in the caller's buffers;
            // its memory stays counted until its scene has been released
            auto publish = [&] {
                {
                    std::lock_guard<std::mutex> lk(m);
                    for (Req *q : g) {
                        q->t_end = Clock::now();
                        q->state = IRTT_REQ_DONE;
                    }
                }
                cv_done.notify_all();
            };
            int rc = g.size() == 1 ? run(*g[0], publish) : run_group(g, publish);
            {
                std::lock_guard<std::mutex> lk(m);
                for (Req *q : g) {
                    in_use -= q->footprint;
                    if (q->state == IRTT_REQ_RUNNING) {
                        q->t_end = Clock::now();
                        q->state = rc == IRTT_OK ? IRTT_REQ_DONE : rc;
                    }
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    static int run(Req &q, const std::function<void()> &publish);
    static int run_group(const std::vector<Req *> &g, const std::function<void()> &publish);
};

namespace {

// scene of a request with its initial state set; nullptr and rc on failure
irtt_scene make_scene(const Req &q, int &rc)
{
    const irtt_request &r = q.r;
    irtt_scene h = nullptr;
    rc = irtt_scene_create(q.xyz.data(), r.nv, q.tri.data(), r.nt, q.mat.data(), r.bin_mode, &h);
    if (rc) return nullptr;
    if (!q.kelvin0.empty() || !q.fixed.empty()) {
        const uint32_t n = r.bin_mode == IRTT_BIN_VERTEX ? r.nv : r.nt;
        std::vector<float> k0 = q.kelvin0.empty() ? std::vector<float>(n, 0.0f) : q.kelvin0;
        std::vector<uint8_t> fx = q.fixed.empty() ? std::vector<uint8_t>(n, 0) : q.fixed;
        rc = r.bin_mode == IRTT_BIN_VERTEX ? irtt_scene_set_vertex_state(h, k0.data(), fx.data())
                                           : irtt_scene_set_bin_state(h, k0.data(), fx.data());
        if (rc) {
            irtt_scene_destroy(h);
            return nullptr;
        }
    }
    return h;
}

// the objective L = 0.5 w sum_i A_i (x_i - target_i)^2 and dL/dx
// (thermal_objective.cpp:34,47-57)
double objective(const Req &q, const double *x, const std::vector<double> &area, double *dLdx)
{
    double L = 0.0;
    for (uint32_t i = 0; i < q.nb; i++) {
        const double d = x[i] - q.target[i];
        L += 0.5 * q.r.objective_weight * area[i] * d * d;
        dLdx[i] = q.r.objective_weight * area[i] * d;
    }
    return L;
}

irtt_adjoint_params adjoint_of(const irtt_request &r)
{
    irtt_adjoint_params ap = r.adjoint;
    if (!(ap.edge_ratio > 0.0f)) ap.edge_ratio = 0.05f;
    if (!(ap.tiny_eps > 0.0f)) ap.tiny_eps = 1e-2f;
    return ap;
}

}  // namespace

// One design: scene -> transport -> solve -> adjoint solve -> adjoint trace
// -> projection (the steps of ThermalRenderer::runForwardSim / runAdjointSim
// for one state, thermal_renderer.cpp:958-1133).
int Server::run(Req &q, const std::function<void()> &publish)
{
    const irtt_request &r = q.r;
    int rc = IRTT_OK;
    irtt_scene h = make_scene(q, rc);
    if (!h) return rc;
    auto done = [&](int code) {
        if (code == IRTT_OK) publish();
        irtt_scene_destroy(h);
        return code;
    };
    const uint32_t nb = q.nb;
    if (!q.source.empty() && (rc = irtt_scene_set_source(h, q.source.data()))) return done(rc);
    irtt_transport_stats ts{};
    // an adjoint trace follows when there is an objective: keep the photons'
    // paths (and absorbing bins) so it needs no replay
    irtt_transport_params tp = r.transport;
    if (!q.target.empty()) tp.flags |= IRTT_TRANSPORT_KEEP_PATHS;
    if ((rc = irtt_build_transport(h, &tp, &ts))) return done(rc);
    if (r.operator_kind != IRTT_OP_DENSE && (rc = irtt_select_operator(h, r.operator_kind)))
        return done(rc);
    irtt_solve_stats ss{};
    if ((rc = irtt_solve(h, &r.solver, &ss))) return done(rc);
    std::vector<double> x(nb);
    if ((rc = irtt_get_temperatures(h, x.data()))) return done(rc);
    if (r.kelvin) std::memcpy(r.kelvin, x.data(), nb * sizeof(double));
    if (q.target.empty()) return done(IRTT_OK);

    std::vector<double> area(nb), dLdx(nb);
    if ((rc = irtt_scene_get_bin_data(h, area.data(), nullptr, nullptr))) return done(rc);
    const double L = objective(q, x.data(), area, dLdx.data());
    if (r.objective) *r.objective = L;
    // adjoint solve, adjoint trace, source term, projection onto the parameters
    const irtt_adjoint_params ap = adjoint_of(r);
    const bool want_dp = r.dp && r.n_params;
    if ((rc = irtt_design_gradient(h, &r.solver, dLdx.data(), &ap, want_dp ? r.nnz : 0,
                                   q.vertex.data(), q.param.data(), q.dxdp.data(),
                                   want_dp ? r.n_params : 0, want_dp ? r.dp : nullptr, nullptr,
                                   nullptr, &ss)))
        return done(rc);
    return done(IRTT_OK);
}

// Designs that share an operator (one group, identical inputs but source and
// target): one scene and transport build; the steady forward solves of all
// of them in lockstep (irtt_solve_batch: batched operator products), then
// the adjoint solves in lockstep (irtt_adjoint_solve_batch), then each
// design's adjoint trace, source term and projection.
int Server::run_group(const std::vector<Req *> &g, const std::function<void()> &publish)
{
    const Req &q0 = *g[0];
    const irtt_request &r0 = q0.r;
    const uint32_t R = (uint32_t)g.size(), nb = q0.nb;
    int rc = IRTT_OK;
    irtt_scene h = make_scene(q0, rc);
    if (!h) return rc;
    auto done = [&](int code) {
        if (code == IRTT_OK) publish();
        irtt_scene_destroy(h);
        return code;
    };
    try {
        Scene &s = *reinterpret_cast<Scene *>(h);
        t_alloc_stream = s.stream;
        irtt_transport_params tp = r0.transport;
        if (!q0.target.empty()) tp.flags |= IRTT_TRANSPORT_KEEP_PATHS;
        irtt_transport_stats ts{};
        if ((rc = irtt_build_transport(h, &tp, &ts))) return done(rc);
        if (r0.operator_kind != IRTT_OP_DENSE && (rc = irtt_select_operator(h, r0.operator_kind)))
            return done(rc);
        t_alloc_stream = s.stream;
        // per-design sources: the request's own, else the materials'
        std::vector<double> src((size_t)R * nb), X((size_t)R * nb);
        for (uint32_t k = 0; k < R; k++) {
            const std::vector<double> &sk = g[k]->source.empty() ? s.bin_source : g[k]->source;
            std::copy(sk.begin(), sk.end(), src.begin() + (size_t)k * nb);
        }
        irtt_solve_stats ss{};
        solve_batch(s, r0.solver, R, src.data(), X.data(), ss);
        for (uint32_t k = 0; k < R; k++)
            if (g[k]->r.kelvin)
                std::memcpy(g[k]->r.kelvin, X.data() + (size_t)k * nb, nb * sizeof(double));
        if (q0.target.empty()) return done(IRTT_OK);
        std::vector<double> rhs((size_t)R * nb), lam((size_t)R * nb);
        for (uint32_t k = 0; k < R; k++) {
            const double L = objective(*g[k], X.data() + (size_t)k * nb, s.bin_area,
                                       rhs.data() + (size_t)k * nb);
            if (g[k]->r.objective) *g[k]->r.objective = L;
        }
        adjoint_solve_batch(s, r0.solver, R, X.data(), rhs.data(), lam.data(), ss);
        for (uint32_t k = 0; k < R; k++) {
            const Req &q = *g[k];
            const bool want_dp = q.r.dp && q.r.n_params;
            // the source term follows the design's own source
            s.source_explicit = !q.source.empty();
            gradient_from_adjoint(s, q.r.solver, X.data() + (size_t)k * nb,
                                  lam.data() + (size_t)k * nb, adjoint_of(q.r),
                                  want_dp ? q.r.nnz : 0, q.vertex.data(), q.param.data(),
                                  q.dxdp.data(), want_dp ? q.r.n_params : 0,
                                  want_dp ? q.r.dp : nullptr, nullptr);
        }
    } catch (const Error &e) {
        set_last_error(e.what());
        return done(e.code);
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return done(IRTT_ERR_STATE);
    }
    return done(IRTT_OK);
}

}  // namespace irtt

using irtt::Error;
using irtt::Req;
using irtt::Server;

namespace {
int sguard(const std::function<void()> &f)
{
    try {
        f();
        return IRTT_OK;
    } catch (const Error &e) {
        irtt::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        irtt::set_last_error("host allocation failed");
        return IRTT_ERR_NOMEM;
    } catch (const std::exception &e) {
        irtt::set_last_error(e.what());
        return IRTT_ERR_STATE;
    }
}

Server *SV(irtt_server h)
{
    if (!h) throw Error(IRTT_ERR_ARG, "null server handle");
    return reinterpret_cast<Server *>(h);
}

void fill_status(const Req &q, irtt_request_status *st)
{
    if (!st) return;
    *st = irtt_request_status{};
    st->state = q.state;
    st->n_bins = q.nb;
    st->footprint_bytes = q.footprint;
    auto now = irtt::Clock::now();
    if (q.state == IRTT_REQ_QUEUED) {
        st->queued_ms = irtt::ms_between(q.t_submit, now);
        return;
    }
    st->queued_ms = irtt::ms_between(q.t_submit, q.t_start);
    auto end = q.state == IRTT_REQ_RUNNING ? now : q.t_end;
    st->run_ms = irtt::ms_between(q.t_start, end);
    st->latency_ms = irtt::ms_between(q.t_submit, end);
}
}  // namespace

extern "C" {

int irtt_serve_create(const irtt_serve_config *cfg, irtt_server *out)
{
    return sguard([&] {
        if (!out) throw Error(IRTT_ERR_ARG, "out is null");
        auto sv = std::make_unique<Server>();
        if (cfg) sv->cfg = *cfg;
        if (!sv->cfg.slots) sv->cfg.slots = 4;
        if (sv->cfg.slots > 64) throw Error(IRTT_ERR_ARG, "at most 64 slots");
        sv->budget = sv->cfg.memory_budget;
        if (!sv->budget) {
            sv->budget = (uint64_t)(0.9 * (double)irtt::dev_available());
        }
        Server *p = sv.get();
        for (uint32_t i = 0; i < sv->cfg.slots; i++) sv->workers.emplace_back([p] { p->worker(); });
        *out = reinterpret_cast<irtt_server>(sv.release());
    });
}

int irtt_serve_submit(irtt_server h, const irtt_request *r, uint64_t *ticket)
{
    return sguard([&] {
        Server *sv = SV(h);
        if (!r || !r->xyz || !r->tri || !r->mat || !r->nv || !r->nt)
            throw Error(IRTT_ERR_ARG, "request without geometry");
        if (r->bin_mode > IRTT_BIN_TRIANGLE) throw Error(IRTT_ERR_ARG, "unknown bin mode");
        if (r->operator_kind != IRTT_OP_DENSE && r->operator_kind != IRTT_OP_DENSE_F16)
            throw Error(IRTT_ERR_ARG, "requests use the dense FP32 or FP16 operator");
        if (r->nnz && (!r->vertex || !r->param || !r->dxdp))
            throw Error(IRTT_ERR_ARG, "parameter map without arrays");
        auto q = std::make_unique<Req>();
        q->r = *r;
        q->nb = r->bin_mode == IRTT_BIN_VERTEX ? r->nv : r->nt;
        q->xyz.assign(r->xyz, r->xyz + 3ull * r->nv);
        q->tri.assign(r->tri, r->tri + 3ull * r->nt);
        q->mat.assign(r->mat, r->mat + r->nt);
        const uint32_t ns = r->bin_mode == IRTT_BIN_VERTEX ? r->nv : r->nt;
        if (r->kelvin0) q->kelvin0.assign(r->kelvin0, r->kelvin0 + ns);
        if (r->fixed) q->fixed.assign(r->fixed, r->fixed + ns);
        if (r->target) q->target.assign(r->target, r->target + q->nb);
        if (r->source) q->source.assign(r->source, r->source + q->nb);
        if (r->nnz) {
            q->vertex.assign(r->vertex, r->vertex + r->nnz);
            q->param.assign(r->param, r->param + r->nnz);
            q->dxdp.assign(r->dxdp, r->dxdp + 3ull * r->nnz);
        }
        q->footprint = irtt::footprint(*r, q->nb);
        if (q->footprint > sv->budget)
            throw Error(IRTT_ERR_LIMIT, "request needs " + std::to_string(q->footprint) +
                                            " bytes of HBM, the budget is " +
                                            std::to_string(sv->budget));
        q->t_submit = irtt::Clock::now();
        {
            std::lock_guard<std::mutex> lk(sv->m);
            if (sv->stop) throw Error(IRTT_ERR_STATE, "server is shutting down");
            q->ticket = sv->next_ticket++;
            if (ticket) *ticket = q->ticket;
            sv->queue.push_back(q.get());
            sv->reqs.emplace(q->ticket, std::move(q));
        }
        sv->cv_work.notify_one();
    });
}

int irtt_serve_poll(irtt_server h, uint64_t ticket, irtt_request_status *st)
{
    return sguard([&] {
        Server *sv = SV(h);
        std::lock_guard<std::mutex> lk(sv->m);
        auto it = sv->reqs.find(ticket);
        if (it == sv->reqs.end()) throw Error(IRTT_ERR_ARG, "unknown ticket");
        fill_status(*it->second, st);
    });
}

int irtt_serve_wait(irtt_server h, uint64_t ticket, irtt_request_status *st)
{
    return sguard([&] {
        Server *sv = SV(h);
        std::unique_lock<std::mutex> lk(sv->m);
        auto it = sv->reqs.find(ticket);
        if (it == sv->reqs.end()) throw Error(IRTT_ERR_ARG, "unknown ticket");
        Req *q = it->second.get();
        sv->cv_done.wait(lk, [&] { return q->state != IRTT_REQ_QUEUED &&
                                          q->state != IRTT_REQ_RUNNING; });
        fill_status(*q, st);
    });
}

int irtt_serve_destroy(irtt_server h)
{
    return sguard([&] {
        if (!h) return;
        Server *sv = SV(h);
        {
            // let the workers drain the queue, then stop
            std::unique_lock<std::mutex> lk(sv->m);
            sv->cv_done.wait(lk, [&] {
                for (auto &kv : sv->reqs)
                    if (kv.second->state == IRTT_REQ_QUEUED || kv.second->state == IRTT_REQ_RUNNING)
                        return false;
                return true;
            });
            sv->stop = true;
        }
        sv->cv_work.notify_all();
        for (auto &t : sv->workers) t.join();
        delete sv;
    });
}

}  // extern "C"
