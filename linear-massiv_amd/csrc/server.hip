// Solve server: continuous admission of variable-size LU / Cholesky solve
// requests (FP32 factor on MFMA, FP64 refinement), an HBM budget, resident
// factors for repeated right-hand sides, and per-request latency.
//
// Requests enter a FIFO queue.  A request is admitted when a worker slot is
// free and its device working set fits in what the budget has left;
// admitted requests run concurrently, each on its worker's own stream, so
// the work of different requests overlaps on the device.  A request that
// does not fit may be overtaken by later ones that do (backfill), unless it
// has waited longer than the aging limit, after which it blocks the queue
// until memory frees up.  Small requests (n <= small_n) are admitted in
// groups and solved together by one batched persistent launch.  A factor
// kept for reuse stays in HBM (its FP64 matrix and FP32 factor) and counts
// against the budget until released.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace lmx {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Factor {
    int64_t id = 0;
    int kind = 0;
    int64_t n = 0;
    int64_t bytes = 0;
    int refs = 0;          // requests using it now
    bool released = false;
    double *A64 = nullptr;  // the FP64 matrix (dense, ld = n)
    void *mem = nullptr;    // the FP32 factor (low_factor_bind)
    bool factored = false;  // small factors are kept as the matrix only
    LowFactor f;
};

struct Request {
    int64_t id = 0;
    int kind = 0;
    int64_t n = 0, nrhs = 0;
    const double *A = nullptr;
    int64_t lda = 0;
    const double *B = nullptr;
    int64_t ldb = 0;
    double *X = nullptr;
    int64_t ldx = 0;
    int64_t factor_in = 0;
    bool keep = false;
    int64_t need = 0;
    bool small = false;
    // a thread is blocked in lmx_server_wait on it (only that thread collects it)
    bool claimed = false;
    // outcome
    bool done = false;
    int rc = 0;
    int iterations = 0;
    int fell_back = 0;
    double eta = 0;
    int64_t factor_out = 0;
    Clock::time_point t_submit, t_start, t_done;
    int ev = -1;  // index of its device interval
};

}  // namespace
}  // namespace lmx

struct lmx_server {
    int device = 0;
    int64_t budget = 0;
    int slots = 4;
    int small_n = 0;
    int small_max = 256;
    double aging_ms = 100.0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<lmx::Request *> pending;
    std::unordered_map<int64_t, lmx::Request *> requests;
    std::unordered_map<int64_t, lmx::Factor *> factors;
    int64_t next_id = 1;
    int64_t used = 0, peak_used = 0;
    int inflight = 0;
    int64_t completed = 0, failed = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    lmx_batch_plan *small_plan = nullptr;
    int64_t small_reserve = 0;
    // device busy intervals
    hipStream_t ref_stream = nullptr;
    hipEvent_t ref = nullptr;
    std::chrono::steady_clock::time_point ref_host;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> intervals;
};

namespace lmx {
namespace {

int64_t req_need(const Request &r, const Factor *reuse) {
    const int64_t n = r.n, k = r.nrhs;
    int64_t b = n * k * (8 + 8 + 8 + 4) + n * 64 * 4 + (4 << 20);  // B, X, R, R32, TRSM blocks, scratch
    if (r.small) return n * n * 12 + b;                                // matrix + FP32 copy, batched
    if (!reuse) b += n * n * 8;                                        // the matrix
    if (!reuse || !reuse->factored) b += (int64_t)low_factor_bytes(n); // an FP32 factor to compute
    b += n * n * 8;                                                    // FP64 fallback copy
    return b;
}
int64_t factor_bytes(int64_t n, bool factored) {
    return n * n * 8 + (factored ? (int64_t)low_factor_bytes(n) : 0);
}

// Under the lock: the next admissible request, or null.
Request *pick(lmx_server *sv, bool small_worker) {
    if (sv->pending.empty()) return nullptr;
    const auto now = Clock::now();
    Request *head = sv->pending.front();
    const bool head_aged = ms_between(head->t_submit, now) > sv->aging_ms;
    for (auto it = sv->pending.begin(); it != sv->pending.end(); ++it) {
        Request *r = *it;
        if (r->small != small_worker) {
            if (r == head && head_aged && r->need > sv->budget - sv->used) return nullptr;
            continue;
        }
        if (r->need <= sv->budget - sv->used) {
            sv->pending.erase(it);
            return r;
        }
        if (r == head && head_aged) return nullptr;  // no overtaking an aged head
    }
    return nullptr;
}

void hip_free_async(void *p, hipStream_t s) {
    if (p) (void)hipFreeAsync(p, s);
}

void copy_in(double *dst, int64_t ldd, const double *src, int64_t lds, int64_t rows, int64_t cols, hipStream_t s) {
    LMX_HIP(hipMemcpy2DAsync(dst, (size_t)ldd * 8, src, (size_t)lds * 8, (size_t)cols * 8, (size_t)rows,
                             hipMemcpyHostToDevice, s));
}
void copy_out(double *dst, int64_t ldd, const double *src, int64_t lds, int64_t rows, int64_t cols, hipStream_t s) {
    LMX_HIP(hipMemcpy2DAsync(dst, (size_t)ldd * 8, src, (size_t)lds * 8, (size_t)cols * 8, (size_t)rows,
                             hipMemcpyDeviceToHost, s));
}

template <class T>
T *dalloc(size_t count, hipStream_t s) {
    void *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMallocAsync(&p, count * sizeof(T), s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(LMX_ERR_NOMEM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes");
    }
    return static_cast<T *>(p);
}

// One large request on the worker's stream.
void run_one(lmx_server *sv, Request *r, Factor *reuse, hipStream_t s) {
    const int64_t n = r->n, k = r->nrhs;
    double *dB = dalloc<double>((size_t)(n * k), s);
    double *dX = dalloc<double>((size_t)(n * k), s);
    try {
        copy_in(dB, k, r->B, r->ldb, n, k, s);
        RefineOut o;
        if (reuse && reuse->factored) {
            o = low_refine(s, reuse->f, k, dB, k, dX, k);
        } else if (reuse) {
            // a resident matrix kept without a factor (it came through the
            // batched path): factor it for this request only
            LowFactor tf;
            tf.spd = reuse->kind == LMX_BATCH_CHOLESKY;
            tf.n = n;
            tf.A = reuse->A64;
            tf.lda = n;
            void *mem = dalloc<char>(low_factor_bytes(n), s);
            low_factor_bind(tf, mem);
            try {
                low_factor(s, tf);
                o = low_refine(s, tf, k, dB, k, dX, k);
            } catch (...) {
                hip_free_async(mem, s);
                throw;
            }
            hip_free_async(mem, s);
        } else {
            // A panel whose workgroups could not all be resident in time
            // (info < 0: a factorisation beside many concurrent requests)
            // is redone once with tournament pivoting, whose panels never
            // wait on each other's workgroups.
            auto factor_refine = [&](LowFactor &lf) {
                low_factor(s, lf);
                if (lf.info < 0 && !lf.spd) {
                    LMX_HIP(hipStreamSynchronize(s));
                    lf.tournament = true;
                    low_factor(s, lf);
                }
                return low_refine(s, lf, k, dB, k, dX, k);
            };
            std::unique_ptr<Factor> f(new Factor());
            f->kind = r->kind;
            f->n = n;
            f->factored = true;
            f->A64 = dalloc<double>((size_t)(n * n), s);
            f->mem = dalloc<char>(low_factor_bytes(n), s);
            f->f.spd = r->kind == LMX_BATCH_CHOLESKY;
            f->f.n = n;
            f->f.A = f->A64;
            f->f.lda = n;
            low_factor_bind(f->f, f->mem);
            try {
                copy_in(f->A64, n, r->A, r->lda, n, n, s);
                o = factor_refine(f->f);
            } catch (...) {
                hip_free_async(f->A64, s);
                hip_free_async(f->mem, s);
                throw;
            }
            if (r->keep) {
                f->bytes = factor_bytes(n, true);
                std::lock_guard<std::mutex> g(sv->mu);
                Factor *fp = f.release();
                fp->id = sv->next_id++;
                r->factor_out = fp->id;
                sv->factors[fp->id] = fp;
            } else {
                hip_free_async(f->A64, s);
                hip_free_async(f->mem, s);
            }
        }
        copy_out(r->X, r->ldx, dX, k, n, k, s);
        LMX_HIP(hipStreamSynchronize(s));
        r->iterations = o.iterations;
        r->fell_back = o.fell_back ? 1 : 0;
        r->eta = o.eta;
        r->rc = o.info > 0 ? o.info : (o.info < 0 ? LMX_ERR_INTERNAL : (o.converged ? LMX_OK : LMX_ERR_NOT_CONVERGED));
    } catch (...) {
        hip_free_async(dB, s);
        hip_free_async(dX, s);
        throw;
    }
    hip_free_async(dB, s);
    hip_free_async(dX, s);
}

// A group of small requests in one batched launch (FP32 factor + FP64
// refinement per problem).
void run_small(lmx_server *sv, std::vector<Request *> &grp, std::vector<Factor *> &reuse, hipStream_t s,
               lmx_batch_plan *plan) {
    std::vector<lmx_batch_problem> probs(grp.size());
    int64_t ao = 0, bo = 0;
    for (size_t i = 0; i < grp.size(); ++i) {
        probs[i] = {grp[i]->kind, (int32_t)grp[i]->n, (int32_t)grp[i]->nrhs, 0, ao, bo, 0};
        ao += grp[i]->n * grp[i]->n;
        bo += grp[i]->n * grp[i]->nrhs;
    }
    double *dA = dalloc<double>((size_t)ao, s);
    double *dB = dalloc<double>((size_t)bo, s);
    int32_t *dI = dalloc<int32_t>(2 * grp.size(), s);
    std::vector<int32_t> hI(2 * grp.size());
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(s);  // no copy into hI may outlive it
        hip_free_async(dA, s);
        hip_free_async(dB, s);
        hip_free_async(dI, s);
    };
    try {
        for (size_t i = 0; i < grp.size(); ++i) {
            Request *r = grp[i];
            const int64_t n = r->n;
            if (reuse[i])
                LMX_HIP(hipMemcpyAsync(dA + probs[i].a_offset, reuse[i]->A64, (size_t)(n * n) * 8,
                                       hipMemcpyDeviceToDevice, s));
            else
                copy_in(dA + probs[i].a_offset, n, r->A, r->lda, n, n, s);
            copy_in(dB + probs[i].b_offset, r->nrhs, r->B, r->ldb, n, r->nrhs, s);
        }
        batch_replan(plan, probs.data(), (int32_t)probs.size(), ao, bo, 0, s);
        const int rc = lmx_batch_plan_execute(plan, dA, dB, nullptr, dI, dI + grp.size(), s);
        if (rc != LMX_OK) throw Error(rc, lmx_last_error());
        LMX_HIP(hipMemcpyAsync(hI.data(), dI, hI.size() * 4, hipMemcpyDeviceToHost, s));
        for (size_t i = 0; i < grp.size(); ++i)
            copy_out(grp[i]->X, grp[i]->ldx, dB + probs[i].b_offset, grp[i]->nrhs, grp[i]->n, grp[i]->nrhs, s);
        // factors kept for reuse: the FP64 matrix stays resident (a small
        // problem is refactored with each reuse, inside the batch)
        std::vector<Factor *> kept(grp.size(), nullptr);
        for (size_t i = 0; i < grp.size(); ++i) {
            if (!grp[i]->keep || reuse[i]) continue;
            std::unique_ptr<Factor> f(new Factor());
            f->kind = grp[i]->kind;
            f->n = grp[i]->n;
            f->A64 = dalloc<double>((size_t)(f->n * f->n), s);
            copy_in(f->A64, f->n, grp[i]->A, grp[i]->lda, f->n, f->n, s);
            f->bytes = factor_bytes(f->n, false);
            kept[i] = f.release();
        }
        LMX_HIP(hipStreamSynchronize(s));
        {
            std::lock_guard<std::mutex> g(sv->mu);
            for (size_t i = 0; i < grp.size(); ++i)
                if (kept[i]) {
                    kept[i]->id = sv->next_id++;
                    grp[i]->factor_out = kept[i]->id;
                    sv->factors[kept[i]->id] = kept[i];
                }
        }
        for (size_t i = 0; i < grp.size(); ++i) {
            const int info = hI[i], it = hI[grp.size() + i];
            grp[i]->rc = info;
            grp[i]->iterations = it < 0 ? 0 : it;
            grp[i]->fell_back = it < 0 ? 1 : 0;
            grp[i]->eta = 0;
        }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

void finish(lmx_server *sv, std::vector<Request *> &rs, std::vector<Factor *> &reuse, int err_code,
            const std::string &err, int ev_index) {
    std::lock_guard<std::mutex> g(sv->mu);
    const auto now = Clock::now();
    for (size_t i = 0; i < rs.size(); ++i) {
        Request *r = rs[i];
        if (err_code != 0) r->rc = err_code;
        r->t_done = now;
        r->done = true;
        r->ev = ev_index;
        sv->used -= r->need;
        if (r->factor_out) sv->used += sv->factors[r->factor_out]->bytes;
        if (reuse[i]) {
            Factor *f = reuse[i];
            if (--f->refs == 0 && f->released) {
                sv->used -= f->bytes;
                sv->factors.erase(f->id);
                hip_free_async(f->A64, sv->ref_stream);
                hip_free_async(f->mem, sv->ref_stream);
                delete f;
            }
        }
        if (r->rc < 0) ++sv->failed;
        ++sv->completed;
    }
    sv->inflight -= (int)rs.size();
    (void)err;
}

void worker_main(lmx_server *sv, bool small_worker, lmx_batch_plan *plan) {
    (void)hipSetDevice(sv->device);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return;
    for (;;) {
        std::vector<Request *> grp;
        std::vector<Factor *> reuse;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] {
                if (sv->stop && sv->pending.empty()) return true;
                Request *r = pick(sv, small_worker);
                if (!r) return false;
                grp.push_back(r);
                return true;
            });
            if (grp.empty()) break;
            if (small_worker)
                while ((int)grp.size() < sv->small_max) {
                    Request *r = pick(sv, true);
                    if (!r) break;
                    grp.push_back(r);
                }
            const auto now = Clock::now();
            for (Request *r : grp) {
                r->t_start = now;
                sv->used += r->need;
                reuse.push_back(r->factor_in ? sv->factors[r->factor_in] : nullptr);
            }
            sv->peak_used = std::max(sv->peak_used, sv->used);
            sv->inflight += (int)grp.size();
        }
        hipEvent_t e0 = nullptr, e1 = nullptr;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e0, s);
        int code = 0;
        std::string msg;
        try {
            if (small_worker)
                run_small(sv, grp, reuse, s, plan);
            else
                run_one(sv, grp[0], reuse[0], s);
        } catch (const Error &e) {
            code = e.code;
            msg = e.what();
        } catch (const std::exception &e) {
            code = LMX_ERR_INTERNAL;
            msg = e.what();
        }
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        int ev_index;
        {
            std::lock_guard<std::mutex> g(sv->mu);
            sv->intervals.emplace_back(e0, e1);
            ev_index = (int)sv->intervals.size() - 1;
        }
        finish(sv, grp, reuse, code, msg, ev_index);
        sv->cv_done.notify_all();
        sv->cv_work.notify_all();
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
}

void fill_result(const Request *r, lmx_request_result *res) {
    if (!res) return;
    res->status = r->done ? 1 : 0;
    res->rc = r->rc;
    res->iterations = r->iterations;
    res->fell_back = r->fell_back;
    res->backward_error = r->eta;
    res->factor = r->factor_out;
    res->queue_ms = r->done || r->t_start != Clock::time_point() ? ms_between(r->t_submit, r->t_start) : 0;
    res->run_ms = r->done ? ms_between(r->t_start, r->t_done) : 0;
    res->latency_ms = r->done ? ms_between(r->t_submit, r->t_done) : 0;
}

}  // namespace
}  // namespace lmx

using namespace lmx;

extern "C" {

int lmx_server_create(lmx_server **out, const lmx_server_config *cfg) {
    return guarded([&]() -> int {
        if (!out) fail_arg("server is NULL");
        *out = nullptr;
        lmx_server_config c{};
        if (cfg) c = *cfg;
        if (c.slots < 0 || c.slots > 16) fail_arg("slots must be in [0, 16] (0 = 4)");
        if (c.small_n < 0 || c.small_n > LMX_BATCH_MAX_N) fail_arg("small_n must be in [0, 512]");
        if (c.budget_bytes < 0) fail_arg("budget_bytes < 0");
        Context &cx = ctx();
        std::unique_ptr<lmx_server> sv(new lmx_server());
        sv->device = cx.device;
        sv->slots = c.slots ? c.slots : 4;
        sv->small_n = c.small_n;
        if (c.small_max > 0) sv->small_max = c.small_max;
        if (c.aging_ms > 0) sv->aging_ms = c.aging_ms;
        size_t fr = 0, tot = 0;
        LMX_HIP(hipMemGetInfo(&fr, &tot));
        sv->budget = c.budget_bytes ? c.budget_bytes : (int64_t)(0.8 * (double)fr);
        LMX_HIP(hipStreamCreateWithFlags(&sv->ref_stream, hipStreamNonBlocking));
        LMX_HIP(hipEventCreate(&sv->ref));
        LMX_HIP(hipEventRecord(sv->ref, sv->ref_stream));
        LMX_HIP(hipEventSynchronize(sv->ref));
        sv->ref_host = Clock::now();
        if (sv->small_n > 0) {
            // the batched worker's plan and its workspace live as long as the
            // server; the workspace is reserved from the budget up front
            lmx_batch_problem one = {LMX_BATCH_LU, sv->small_n, 1, 0, 0, 0, 0};
            int rc = lmx_batch_plan_create(&sv->small_plan, &one, 1, (int64_t)sv->small_n * sv->small_n,
                                           sv->small_n, 0, LMX_PREC_F32, 0);
            if (rc != LMX_OK) throw Error(rc, lmx_last_error());
            sv->small_reserve = (int64_t)batch_slot_bytes(sv->small_n, LMX_PREC_F32, 0) *
                                std::min(sv->small_max, 2 * cx.cu_count);
            if (sv->small_reserve >= sv->budget) {
                lmx_batch_plan_destroy(sv->small_plan);
                fail_arg("budget too small for the batched worker's workspace");
            }
            sv->used = sv->small_reserve;
        }
        lmx_server *p = sv.get();
        for (int i = 0; i < sv->slots; ++i) sv->workers.emplace_back(worker_main, p, false, nullptr);
        if (sv->small_n > 0) sv->workers.emplace_back(worker_main, p, true, sv->small_plan);
        *out = sv.release();
        return LMX_OK;
    });
}

int lmx_server_submit(lmx_server *sv, int32_t kind, int64_t n, int64_t nrhs, const double *A, int64_t lda,
                      int64_t factor, const double *B, int64_t ldb, double *X, int64_t ldx, int32_t flags,
                      int64_t *id) {
    return guarded([&]() -> int {
        if (!sv || !id) fail_arg("server and id must not be NULL");
        if (kind != LMX_BATCH_LU && kind != LMX_BATCH_CHOLESKY) fail_arg("kind must be LU or Cholesky");
        if (n < 1 || nrhs < 1) fail_arg("n and nrhs must be >= 1");
        if (!B || !X) fail_arg("B and X must not be NULL");
        if (ldb < nrhs || ldx < nrhs) fail_arg("bad ldb / ldx");
        if (flags & ~LMX_REQ_KEEP_FACTOR) fail_arg("unknown request flags");
        if (factor == 0 && (!A || lda < n)) fail_arg("A (with lda >= n) or a factor handle is required");
        std::unique_ptr<Request> r(new Request());
        r->kind = kind;
        r->n = n;
        r->nrhs = nrhs;
        r->A = A;
        r->lda = lda;
        r->B = B;
        r->ldb = ldb;
        r->X = X;
        r->ldx = ldx;
        r->factor_in = factor;
        r->keep = (flags & LMX_REQ_KEEP_FACTOR) != 0 && factor == 0;
        r->t_submit = Clock::now();
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->stop) fail_arg("server is shutting down");
        Factor *f = nullptr;
        if (factor) {
            auto it = sv->factors.find(factor);
            if (it == sv->factors.end() || it->second->released) fail_arg("unknown or released factor handle");
            f = it->second;
            if (f->kind != kind || f->n != n) fail_arg("factor handle does not match kind / n");
        }
        r->small = sv->small_n > 0 && n <= sv->small_n && nrhs <= LMX_BATCH_MAX_NRHS && (!f || !f->factored);
        r->need = req_need(*r, f);
        const int64_t resident = r->keep ? factor_bytes(n, !r->small) : 0;
        if (r->need + resident > sv->budget)
            throw Error(LMX_ERR_BUDGET, "request needs " + std::to_string(r->need + resident) +
                                            " bytes of HBM, more than the server budget " +
                                            std::to_string(sv->budget));
        r->id = sv->next_id++;
        *id = r->id;
        if (f) ++f->refs;  // held from submission to completion
        sv->requests[r->id] = r.get();
        sv->pending.push_back(r.release());
        sv->cv_work.notify_all();
        return LMX_OK;
    });
}

int lmx_server_poll(lmx_server *sv, int64_t id, lmx_request_result *res) {
    return guarded([&]() -> int {
        if (!sv) fail_arg("server is NULL");
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->requests.find(id);
        if (it == sv->requests.end()) fail_arg("unknown request id");
        Request *r = it->second;
        if (r->claimed) return 1;  // another thread waits for it and will collect it
        fill_result(r, res);
        if (!r->done) return 1;
        sv->requests.erase(it);
        delete r;
        return LMX_OK;
    });
}

int lmx_server_wait(lmx_server *sv, int64_t id, lmx_request_result *res) {
    return guarded([&]() -> int {
        if (!sv) fail_arg("server is NULL");
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->requests.find(id);
        if (it == sv->requests.end()) fail_arg("unknown request id");
        Request *r = it->second;
        if (r->claimed) fail_arg("another thread is already waiting for this request");
        r->claimed = true;  // from here on only this thread erases and deletes r
        sv->cv_done.wait(lk, [&] { return r->done; });
        fill_result(r, res);
        sv->requests.erase(id);
        delete r;
        return LMX_OK;
    });
}

int lmx_server_factor_release(lmx_server *sv, int64_t factor) {
    return guarded([&]() -> int {
        if (!sv) fail_arg("server is NULL");
        std::lock_guard<std::mutex> g(sv->mu);
        auto it = sv->factors.find(factor);
        if (it == sv->factors.end() || it->second->released) fail_arg("unknown or released factor handle");
        Factor *f = it->second;
        f->released = true;  // freed now, or when its last request completes
        if (f->refs == 0) {
            sv->used -= f->bytes;
            sv->factors.erase(it);
            hip_free_async(f->A64, sv->ref_stream);
            hip_free_async(f->mem, sv->ref_stream);
            delete f;
        }
        sv->cv_work.notify_all();
        return LMX_OK;
    });
}

int lmx_server_stats(lmx_server *sv, lmx_server_counters *st, int32_t reset) {
    return guarded([&]() -> int {
        if (!sv || !st) fail_arg("server and stats must not be NULL");
        std::vector<std::pair<hipEvent_t, hipEvent_t>> iv;
        Clock::time_point ref_host;
        hipEvent_t ref;
        {
            std::lock_guard<std::mutex> g(sv->mu);
            st->completed = sv->completed;
            st->failed = sv->failed;
            st->queued = (int64_t)sv->pending.size();
            st->inflight = sv->inflight;
            st->bytes_in_use = sv->used;
            st->peak_bytes = sv->peak_used;
            st->budget_bytes = sv->budget;
            st->resident_factors = (int64_t)sv->factors.size();
            iv = sv->intervals;
            ref = sv->ref;
            ref_host = sv->ref_host;
            if (reset) {
                sv->intervals.clear();
                sv->peak_used = sv->used;
            }
        }
        // union of the device intervals since the last reset, over the
        // wall time since then
        std::vector<std::pair<double, double>> spans;
        spans.reserve(iv.size());
        for (auto &p : iv) {
            float a = 0, b = 0;
            if (hipEventElapsedTime(&a, ref, p.first) == hipSuccess &&
                hipEventElapsedTime(&b, ref, p.second) == hipSuccess)
                spans.emplace_back(a, b);
        }
        std::sort(spans.begin(), spans.end());
        double busy = 0, cur_a = -1, cur_b = -1;
        for (auto &sp : spans) {
            if (sp.first > cur_b) {
                if (cur_b > cur_a) busy += cur_b - cur_a;
                cur_a = sp.first;
                cur_b = sp.second;
            } else {
                cur_b = std::max(cur_b, sp.second);
            }
        }
        if (cur_b > cur_a) busy += cur_b - cur_a;
        const double wall = ms_between(ref_host, Clock::now());
        st->busy_ms = busy;
        st->window_ms = wall;
        if (reset) {
            for (auto &p : iv) {
                (void)hipEventDestroy(p.first);
                (void)hipEventDestroy(p.second);
            }
            std::lock_guard<std::mutex> g(sv->mu);
            LMX_HIP(hipEventRecord(sv->ref, sv->ref_stream));
            LMX_HIP(hipEventSynchronize(sv->ref));
            sv->ref_host = Clock::now();
        }
        return LMX_OK;
    });
}

int lmx_server_destroy(lmx_server *sv) {
    return guarded([&]() -> int {
        if (!sv) return LMX_OK;
        {
            std::lock_guard<std::mutex> g(sv->mu);
            sv->stop = true;
            // queued requests are not started
            for (Request *r : sv->pending) {
                r->rc = LMX_ERR_DEVICE;
                r->done = true;
            }
            sv->pending.clear();
        }
        sv->cv_work.notify_all();
        sv->cv_done.notify_all();
        for (auto &t : sv->workers) t.join();
        if (sv->small_plan) lmx_batch_plan_destroy(sv->small_plan);
        for (auto &kv : sv->requests) delete kv.second;
        for (auto &kv : sv->factors) {
            hip_free_async(kv.second->A64, sv->ref_stream);
            hip_free_async(kv.second->mem, sv->ref_stream);
            delete kv.second;
        }
        for (auto &p : sv->intervals) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        (void)hipStreamSynchronize(sv->ref_stream);
        (void)hipEventDestroy(sv->ref);
        (void)hipStreamDestroy(sv->ref_stream);
        delete sv;
        return LMX_OK;
    });
}

}  // extern "C"
