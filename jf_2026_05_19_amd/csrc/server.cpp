// server.cpp -- device-resident SOL 105 request server (include/sol105.h,
// sol105_server_*).
//
// The reference serves design iterations through a JSONL worker that runs
// one manifest case after another on the CPU (JFEM/tools/jfem_worker_jsonl.jl,
// JFEM/tools/manifest_batch_core.jl:409-478).  Here a client registers each
// mesh once (uploaded and analysed once) and then submits designs of it:
// new shell thicknesses only, so the pattern, ordering, solve plans and
// factor buffers stay on the device between design iterations.  Worker
// threads, each with its own HIP stream, take requests as soon as a worker
// and device memory are free, so solves of different requests overlap on
// the device.  Requests of small meshes that are pending together are
// stacked into one batched solve (sol105_batch_*).  A request is admitted
// when its estimated device footprint fits the budget beside the work in
// flight.  On request (SOL105_SERVER_WARM_START) each worker context starts
// a mesh's eigensolve from the Ritz vectors of that mesh's previous design
// (SOL105_OPT_WARM_START).  A request that repeats a mesh's last solved
// design is answered from that solve's results.
#include "internal.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

using namespace s105;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes a solve of M holds: model, element blocks, K and KG, fronts,
// explicit inverses, solve workspace and the Lanczos basis; before the first
// analysis an estimate per DOF (measured on the panel workloads: ~9 kB of
// fronts per DOF, the same again for the inverses and the basis).
double model_footprint(const sol105_model *M) {
    const double ndof = M->ndof();
    if (!M->have_sym) return 40e3 * ndof;
    const double nf = M->sym.n_free;
    double b = 8.0 * (double)M->sym.front_storage + 16.0 * (double)(M->splan.p_storage + M->splan.p_pad);
    b += 8.0 * 36.0 * 2.0 * (double)M->pat.n_blocks();
    b += 8.0 * 36.0 * (16.0 * M->n_q4 + 9.0 * M->n_t3 + 4.0 * M->n_bar);
    b += 8.0 * (double)solve_w_size(M->sym.vec_storage);
    b += 8.0 * nf * std::min(1040.0, nf + 16.0) + 8.0 * 4.0 * ndof * 16.0;
    return b;
}

struct SrvModel {
    std::unique_ptr<sol105_model> M;
    bool busy = false;
    int nd = 3;
    // the last design solved alone and its results (a request that repeats
    // it is answered without a solve)
    bool last_valid = false;
    std::vector<double> last_t, last_lambda, last_modes;  // last_modes empty when not kept
    int last_max_modes = 0;
};

struct Request {
    int model = 0;
    std::vector<double> t;
    int max_modes = 16;
    Clock::time_point submitted, started, finished;
    bool done = false;
    int status = SOL105_OK;
    int members = 1;
    double lanczos_vectors = 0.0;  // basis vectors its eigensolve generated
    bool cached = false;           // answered from the last solve of the same design
    bool want_modes = false;       // keep the mode shapes (SOL105_REQ_MODES)
    std::vector<double> lambda;
    std::vector<double> modes;     // n_out x 6 n_nodes, mode-major
};

// a request and its mesh, both resolved under the server lock
struct Job {
    Request *r;
    SrvModel *m;
};

}  // namespace

struct sol105_server {
    int device = 0;
    double budget = 0.0, inflight = 0.0;
    int small_dof = 30000, max_members = 64;
    bool batching = true;
    sol105_ctx *ctx0 = nullptr;  // registration (uploads) on the client's thread
    std::deque<std::unique_ptr<SrvModel>> models;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, Request> reqs;
    int64_t next_ticket = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::thread> workers;
    bool stop = false;
    Clock::time_point t0 = Clock::now();
    std::string err;
    int64_t n_done = 0, n_batches = 0, n_batched = 0, n_cached = 0;
    double peak_inflight = 0.0;
    bool warm_start = false;
};

namespace {

int srv_err(sol105_server *S, int code, const std::string &m) {
    if (S) S->err = m;
    return code;
}

// Picks the next group under the lock: the first admissible request in
// arrival order, plus (when it is small and batching is on) the other
// pending small requests of idle meshes that fit the budget.
std::vector<int64_t> pick(sol105_server *S, double &fp) {
    std::vector<int64_t> group;
    std::vector<int> used;
    fp = 0.0;
    bool small = false;
    const sol105_model *lead = nullptr;  // stacked members share its K6ROT / AUTOSPC
    for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
        Request &R = S->reqs[*it];
        SrvModel &SM = *S->models[R.model];
        if (SM.busy || std::find(used.begin(), used.end(), R.model) != used.end()) continue;
        double f = model_footprint(SM.M.get());
        if (group.empty()) {
            if (S->inflight > 0.0 && S->inflight + f > S->budget) continue;
            group.push_back(*it);
            used.push_back(R.model);
            fp = f;
            lead = SM.M.get();
            small = S->batching && SM.M->ndof() <= S->small_dof && !SM.M->has_mpc();
            if (!small) break;
        } else if (SM.M->ndof() <= S->small_dof && !SM.M->has_mpc() && (int)group.size() < S->max_members &&
                   S->inflight + fp + f <= S->budget && SM.M->k6rot == lead->k6rot &&
                   (SM.M->autospc != 0) == (lead->autospc != 0)) {
            group.push_back(*it);
            used.push_back(R.model);
            fp += f;
        }
    }
    return group;
}

// The design a request asks for: its own thicknesses, or the mesh's current
// ones when it gave none.
std::vector<double> design_of(const Request &r, const sol105_model *M) {
    if (!r.t.empty()) return r.t;
    std::vector<double> t(M->n_q4 + M->n_t3);
    for (int e = 0; e < M->n_q4; e++) t[e] = M->q4prop[6 * (size_t)e + 2];
    for (int e = 0; e < M->n_t3; e++) t[M->n_q4 + e] = M->t3prop[6 * (size_t)e + 2];
    return t;
}

// Runs a group outside the server lock.  Requests and meshes of the group
// belong to this worker until it marks them done (their meshes are busy), so
// no other thread touches them meanwhile.
void run_group(sol105_ctx *ctx, const std::vector<Job> &jobs) {
    for (const Job &j : jobs) j.r->started = Clock::now();
    if (jobs.size() == 1) {
        Request &r = *jobs[0].r;
        SrvModel &SM = *jobs[0].m;
        sol105_model *M = SM.M.get();
        const std::vector<double> t = design_of(r, M);
        r.members = 1;
        const size_t ndof = (size_t)M->ndof();
        if (SM.last_valid && t == SM.last_t && r.max_modes <= SM.last_max_modes &&
            (!r.want_modes || !SM.last_modes.empty())) {
            const size_t k = std::min<size_t>(SM.last_lambda.size(), (size_t)r.max_modes);
            r.status = SOL105_OK;
            r.cached = true;
            r.lambda.assign(SM.last_lambda.begin(), SM.last_lambda.begin() + k);
            if (r.want_modes) r.modes.assign(SM.last_modes.begin(), SM.last_modes.begin() + k * ndof);
        } else {
            int rc = SOL105_OK;
            if (!r.t.empty()) rc = sol105_model_set_shell_thickness(ctx, M, r.t.data());
            std::vector<double> lam(std::max(r.max_modes, 1));
            std::vector<double> modes(r.want_modes ? (size_t)std::max(r.max_modes, 1) * ndof : 0);
            int32_t n = 0;
            if (rc == SOL105_OK)
                rc = sol105_solve(ctx, M, r.max_modes, lam.data(), r.want_modes ? modes.data() : nullptr, &n);
            r.status = rc;
            SM.last_valid = rc == SOL105_OK;
            if (rc == SOL105_OK) {
                r.lambda.assign(lam.begin(), lam.begin() + n);
                if (r.want_modes) modes.resize((size_t)n * ndof);
                r.modes = modes;
                r.lanczos_vectors = M->lanczos_m;
                SM.last_t = t;
                SM.last_lambda = r.lambda;
                SM.last_modes = std::move(modes);
                SM.last_max_modes = r.max_modes;
            }
        }
    } else {
        // stacked solve of the group: member descriptions point into the
        // registered models' host arrays with the requested thicknesses
        std::vector<sol105_model_desc> d(jobs.size());
        int mm = 1;
        for (size_t i = 0; i < jobs.size(); i++) {
            sol105_model *M = jobs[i].m->M.get();
            const Request &r = *jobs[i].r;
            // the registered model takes the design too, so a later solve of
            // it alone starts from the same thicknesses
            if (!r.t.empty()) (void)sol105_model_set_shell_thickness(ctx, M, r.t.data());
            jobs[i].m->last_valid = false;
            sol105_model_desc &x = d[i];
            x.n_nodes = M->n_nodes;
            x.xyz = M->xyz.data();
            x.n_q4 = M->n_q4;
            x.q4_nodes = M->q4.data();
            x.q4_prop = M->q4prop.data();
            x.n_t3 = M->n_t3;
            x.t3_nodes = M->t3.data();
            x.t3_prop = M->t3prop.data();
            x.n_bar = M->n_bar;
            x.bar_nodes = M->bar.data();
            x.bar_vref = M->bar_vref.data();
            x.bar_prop = M->barprop.data();
            x.spc_mask = M->spc_mask.data();
            x.force = M->force.data();
            x.node_has_shell = M->has_shell.data();
            x.node_has_bar = M->has_bar.data();
            x.k6rot = M->k6rot;
            x.autospc = M->autospc;
            x.nd = M->nd;
            x.eigrl_v1 = M->v1;
            x.eigrl_v2 = M->v2;
            x.n_shell_cards = M->n_shell_cards;
            mm = std::max(mm, r.max_modes);
        }
        sol105_batch *B = nullptr;
        int rc = sol105_batch_create(ctx, d.data(), (int32_t)d.size(), &B);
        std::vector<double> lam((size_t)mm * jobs.size());
        std::vector<int32_t> n(jobs.size()), st(jobs.size());
        if (rc == SOL105_OK) rc = sol105_batch_solve(ctx, B, mm, lam.data(), n.data(), st.data());
        double bst[8] = {0};
        if (rc == SOL105_OK) (void)sol105_batch_stats(ctx, B, bst);
        for (size_t i = 0; i < jobs.size(); i++) {
            Request &r = *jobs[i].r;
            r.members = (int)jobs.size();
            r.status = rc != SOL105_OK ? rc : st[i];
            r.lanczos_vectors = bst[1];
            if (r.status == SOL105_OK) {
                int k = std::min<int>(n[i], r.max_modes);
                r.lambda.assign(lam.begin() + (size_t)i * mm, lam.begin() + (size_t)i * mm + k);
                if (r.want_modes) {
                    const size_t ndof = (size_t)jobs[i].m->M->ndof();
                    std::vector<double> all((size_t)std::max(n[i], 1) * ndof);
                    if (sol105_batch_modes(ctx, B, (int32_t)i, all.data()) == SOL105_OK) {
                        all.resize((size_t)k * ndof);
                        r.modes = std::move(all);
                    } else {
                        r.status = SOL105_E_DEVICE;
                    }
                }
            }
        }
        if (B) sol105_batch_free(ctx, B);
    }
    Clock::time_point fin = Clock::now();
    for (const Job &j : jobs) j.r->finished = fin;
}

void worker(sol105_server *S) {
    sol105_ctx *ctx = nullptr;
    if (sol105_ctx_create(S->device, &ctx) != SOL105_OK) return;
    if (S->warm_start) (void)sol105_ctx_set_option(ctx, SOL105_OPT_WARM_START, 1);
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        double fp = 0.0;
        std::vector<int64_t> group;
        S->cv_work.wait(lk, [&] {
            if (S->stop && S->queue.empty()) return true;
            group = pick(S, fp);
            return !group.empty();
        });
        if (group.empty()) break;  // stopping with an empty queue
        std::vector<Job> jobs;
        for (int64_t tk : group) {
            S->queue.erase(std::find(S->queue.begin(), S->queue.end(), tk));
            Request *r = &S->reqs.at(tk);  // element addresses survive rehashing
            SrvModel *m = S->models[r->model].get();
            m->busy = true;
            jobs.push_back({r, m});
        }
        S->inflight += fp;
        S->peak_inflight = std::max(S->peak_inflight, S->inflight);
        lk.unlock();
        run_group(ctx, jobs);
        lk.lock();
        S->inflight -= fp;
        if (S->inflight < 1.0) S->inflight = 0.0;
        for (const Job &j : jobs) {
            j.r->done = true;
            j.m->busy = false;
            S->n_cached += j.r->cached;
        }
        S->n_done += (int64_t)group.size();
        if (group.size() > 1) {
            S->n_batches++;
            S->n_batched += (int64_t)group.size();
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
    lk.unlock();
    sol105_ctx_destroy(ctx);
}

void fill_result(const Request &r, const sol105_server *S, double *lambda_out, int32_t *n_out, int32_t *status,
                 double *info, double *modes_out = nullptr) {
    if (modes_out)
        for (size_t t = 0; t < r.modes.size(); t++) modes_out[t] = r.modes[t];
    if (n_out) *n_out = (int32_t)r.lambda.size();
    if (status) *status = r.status;
    if (lambda_out)
        for (size_t t = 0; t < r.lambda.size(); t++) lambda_out[t] = r.lambda[t];
    if (info) {
        info[0] = ms_since(r.submitted, r.finished);
        info[1] = ms_since(r.submitted, r.started);
        info[2] = ms_since(r.started, r.finished);
        info[3] = r.members;
        info[4] = ms_since(S->t0, r.started);
        info[5] = ms_since(S->t0, r.finished);
        info[6] = r.lanczos_vectors;
        info[7] = r.cached ? 1.0 : 0.0;
    }
}

}  // namespace

extern "C" {

int sol105_server_create(int device, int32_t n_workers, double hbm_budget_bytes, int32_t flags,
                         sol105_server **out) {
    if (!out || n_workers <= 0) return SOL105_E_INVALID;
    *out = nullptr;
    auto S = std::make_unique<sol105_server>();
    S->device = device;
    if (sol105_ctx_create(device, &S->ctx0) != SOL105_OK) return SOL105_E_DEVICE;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    S->budget = hbm_budget_bytes > 0.0 ? hbm_budget_bytes : 0.75 * (double)free_b;
    S->batching = (flags & SOL105_SERVER_STACKING) != 0;
    S->warm_start = (flags & SOL105_SERVER_WARM_START) != 0;
    for (int w = 0; w < n_workers; w++) S->workers.emplace_back(worker, S.get());
    *out = S.release();
    return SOL105_OK;
}

int sol105_server_destroy(sol105_server *S) {
    if (!S) return SOL105_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    S->models.clear();
    sol105_ctx_destroy(S->ctx0);
    delete S;
    return SOL105_OK;
}

const char *sol105_server_last_error(const sol105_server *S) { return S ? S->err.c_str() : "null server"; }

int sol105_server_add_model(sol105_server *S, const sol105_model_desc *desc, int32_t *model_id) {
    if (!S || !desc || !model_id) return srv_err(S, SOL105_E_INVALID, "null argument");
    auto SM = std::make_unique<SrvModel>();
    SM->M = std::make_unique<sol105_model>();
    int rc = model_init(S->ctx0, desc, SM->M.get());
    if (rc) return srv_err(S, rc, sol105_last_error(S->ctx0));
    SM->nd = desc->nd;
    std::lock_guard<std::mutex> lk(S->mu);
    S->models.push_back(std::move(SM));
    *model_id = (int32_t)S->models.size() - 1;
    return SOL105_OK;
}

int sol105_server_submit(sol105_server *S, int32_t model_id, const double *shell_thickness, int32_t max_modes,
                         int64_t *ticket) {
    return sol105_server_submit_ex(S, model_id, shell_thickness, max_modes, 0, ticket);
}

int sol105_server_submit_ex(sol105_server *S, int32_t model_id, const double *shell_thickness, int32_t max_modes,
                            int32_t flags, int64_t *ticket) {
    if (!S || !ticket || max_modes <= 0) return srv_err(S, SOL105_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (model_id < 0 || model_id >= (int32_t)S->models.size())
        return srv_err(S, SOL105_E_INVALID, "unknown model id");
    const sol105_model *M = S->models[model_id]->M.get();
    Request r;
    r.model = model_id;
    r.max_modes = std::min<int32_t>(max_modes, SYMEIG_MAX_K);
    r.want_modes = (flags & SOL105_REQ_MODES) != 0;
    if (shell_thickness) {
        r.t.assign(shell_thickness, shell_thickness + M->n_q4 + M->n_t3);
        for (double v : r.t)
            if (!(v > 0.0)) return srv_err(S, SOL105_E_INVALID, "thickness must be positive");
    }
    r.submitted = Clock::now();
    const int64_t tk = S->next_ticket++;
    S->reqs.emplace(tk, std::move(r));
    S->queue.push_back(tk);
    *ticket = tk;
    S->cv_work.notify_one();
    return SOL105_OK;
}

int sol105_server_poll(sol105_server *S, int64_t ticket, double *lambda_out, int32_t *n_out, int32_t *status,
                       double *info) {
    if (!S) return SOL105_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return srv_err(S, SOL105_E_INVALID, "unknown ticket");
    if (!it->second.done) return SOL105_E_PENDING;
    fill_result(it->second, S, lambda_out, n_out, status, info);
    S->reqs.erase(it);
    S->cv_done.notify_all();  // a waiter of this ticket sees it released
    return SOL105_OK;
}

int sol105_server_wait(sol105_server *S, int64_t ticket, double *lambda_out, int32_t *n_out, int32_t *status,
                       double *info) {
    return sol105_server_wait_modes(S, ticket, lambda_out, n_out, status, info, nullptr);
}

int sol105_server_wait_modes(sol105_server *S, int64_t ticket, double *lambda_out, int32_t *n_out, int32_t *status,
                             double *info, double *modes_out) {
    if (!S) return SOL105_E_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->reqs.find(ticket) == S->reqs.end()) return srv_err(S, SOL105_E_INVALID, "unknown ticket");
    // look the ticket up again after every wake-up: submits may rehash the
    // map, and a concurrent poll or wait of the same ticket may release it
    bool released = false;
    S->cv_done.wait(lk, [&] {
        auto it = S->reqs.find(ticket);
        released = it == S->reqs.end();
        return released || it->second.done;
    });
    if (released) return srv_err(S, SOL105_E_INVALID, "ticket released by another caller");
    auto it = S->reqs.find(ticket);
    fill_result(it->second, S, lambda_out, n_out, status, info, modes_out);
    S->reqs.erase(it);
    S->cv_done.notify_all();  // other waiters of this ticket see it released
    return SOL105_OK;
}

int sol105_server_stats(sol105_server *S, double *out) {
    if (!S || !out) return SOL105_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    out[0] = (double)S->n_done;
    out[1] = (double)S->n_batches;
    out[2] = (double)S->n_batched;
    out[3] = (double)S->queue.size();
    out[4] = S->inflight;
    out[5] = S->budget;
    out[6] = (double)S->workers.size();
    out[7] = (double)S->models.size();
    out[8] = (double)S->n_cached;
    out[9] = S->peak_inflight;
    return SOL105_OK;
}

}  // extern "C"
