// serve.hip -- request loop over device-resident matrices with continuous
// batching (host code only).
//
// The reference answers one query at a time: reset, upload the query, run,
// read back (host_spmv_ucsr.cpp:1084-1180; cpu_mkl_baseline.cpp:254-272).
// Here a worker thread keeps every registered matrix resident in device
// memory and turns the queue of arriving requests into batches:
//   * batches run concurrently on separate streams ("slots"), each slot with
//     a batch workspace of its own, so a matrix may have as many batches in
//     flight as there are slots;
//   * whenever a slot is free, the matrix whose oldest waiting request is the
//     oldest gets the slot and up to max_batch of its waiting requests, each
//     with its own K -- requests that arrive while batches run are admitted
//     into the next one as soon as a slot frees (continuous batching);
//   * matrices are admitted against a device-memory budget that includes
//     every slot's workspace for a full batch of the matrix; requests against
//     a host-memory budget (the query, its results and bookkeeping) and the
//     queue capacity.
// Results are those of acc_gpu_topk_lowbit_batch, which do not depend on the
// batch a query lands in.
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

using acc::fail;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int kMaxSlots = 8;
constexpr int kMaxK = 960;

struct Request {
    int64_t bytes = 0;  // charged to the request-memory budget
    int matrix = 0;
    int k = 0;
    bool keep = true;  // false: results are discarded (load generator)
    std::vector<float> x;
    std::vector<acc_element> out;
    Clock::time_point t0, t1;
    bool done = false;
    int status = ACC_OK;
};

struct Matrix {
    acc_gpu_matrix *m = nullptr;
    int cols = 0;
    std::deque<uint64_t> pending;
    int inflight = 0;   // batches on the slots
    int64_t bytes = 0;
};

struct Slot {
    acc::BatchWorkspace *ws = nullptr;  // this slot's batch workspace (any matrix)
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    float *x_pin = nullptr;        // max_batch x max_cols
    acc_element *o_pin = nullptr;  // max_batch x kMaxK
    float *x_dev = nullptr;
    acc_element *o_dev = nullptr;
    std::vector<uint64_t> ids;
    std::vector<int32_t> ks;
    int matrix = -1;
    int stride = 0;
    bool busy = false;
};

}  // namespace

struct acc_server {
    acc_server_opts o{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<Matrix> mats;
    Slot slots[kMaxSlots];
    std::thread worker;
    bool stop = false;
    uint64_t next_id = 1;
    acc::Budget mem;  // device bytes: slot buffers and resident matrices with their workspaces
    int max_cols = 8192;
    int64_t npending = 0;
    int64_t ws_need = 0;    // the largest full-batch workspace of a resident matrix (reserved per slot)
    int64_t req_bytes = 0;  // host bytes held by waiting and unread requests
    // statistics since the last reset
    std::vector<double> lat_us;
    int64_t submitted = 0, completed = 0, refused = 0, batches = 0, batched = 0;
    int64_t peak_inflight = 0;  // most batches of one matrix on the slots at once
    Clock::time_point t_reset = Clock::now();
};

namespace {

// Forgets a request and returns its bytes to the request-memory budget.
void erase_request(acc_server *s, std::unordered_map<uint64_t, Request>::iterator it) {
    s->req_bytes -= it->second.bytes;
    s->reqs.erase(it);
}

// Completed batches: copy results out, release the slot.
void reap(acc_server *s) {
    for (int i = 0; i < s->o.slots; ++i) {
        Slot &sl = s->slots[i];
        if (!sl.busy) continue;
        const hipError_t e = hipEventQuery(sl.ev);
        if (e == hipErrorNotReady) continue;
        const int status = e == hipSuccess ? ACC_OK : ACC_ERR_DEVICE;
        const auto now = Clock::now();
        for (size_t j = 0; j < sl.ids.size(); ++j) {
            auto it = s->reqs.find(sl.ids[j]);
            if (it == s->reqs.end()) continue;
            Request &r = it->second;
            if (r.keep && status == ACC_OK) r.out.assign(sl.o_pin + j * sl.stride, sl.o_pin + j * sl.stride + r.k);
            r.t1 = now;
            r.done = true;
            r.status = status;
            s->lat_us.push_back(std::chrono::duration<double, std::micro>(r.t1 - r.t0).count());
            ++s->completed;
            if (!r.keep) erase_request(s, it);
        }
        --s->mats[(size_t)sl.matrix].inflight;
        sl.busy = false;
        sl.ids.clear();
    }
}

// Gives free slots to the matrices whose oldest waiting request is oldest.
void launch(acc_server *s) {
    for (int i = 0; i < s->o.slots; ++i) {
        Slot &sl = s->slots[i];
        if (sl.busy) continue;
        int best = -1;
        Clock::time_point oldest = Clock::time_point::max();
        for (size_t mi = 0; mi < s->mats.size(); ++mi) {
            Matrix &mx = s->mats[mi];
            if (!mx.m || mx.pending.empty()) continue;
            const Clock::time_point t = s->reqs[mx.pending.front()].t0;
            if (t < oldest) {
                oldest = t;
                best = (int)mi;
            }
        }
        if (best < 0) return;
        Matrix &mx = s->mats[(size_t)best];
        const int nq = (int)std::min<size_t>(mx.pending.size(), (size_t)s->o.max_batch);
        sl.ids.assign(mx.pending.begin(), mx.pending.begin() + nq);
        mx.pending.erase(mx.pending.begin(), mx.pending.begin() + nq);
        s->npending -= nq;
        sl.ks.resize((size_t)nq);
        int kmax = 1;
        for (int j = 0; j < nq; ++j) {
            const Request &r = s->reqs[sl.ids[(size_t)j]];
            std::memcpy(sl.x_pin + (size_t)j * mx.cols, r.x.data(), (size_t)mx.cols * 4);
            sl.ks[(size_t)j] = r.k;
            kmax = std::max(kmax, r.k);
        }
        sl.stride = kmax;
        sl.matrix = best;
        int rc = hipMemcpyAsync(sl.x_dev, sl.x_pin, (size_t)nq * mx.cols * 4, hipMemcpyHostToDevice, sl.st) == hipSuccess
                     ? ACC_OK : ACC_ERR_DEVICE;
        if (rc == ACC_OK)
            rc = acc::topk_lowbit_batch_ws(mx.m, sl.ws, sl.x_dev, nq, sl.ks.data(), s->o.ck, sl.o_dev, kmax, sl.st);
        if (rc == ACC_OK &&
            hipMemcpyAsync(sl.o_pin, sl.o_dev, (size_t)nq * kmax * sizeof(acc_element), hipMemcpyDeviceToHost, sl.st) != hipSuccess)
            rc = ACC_ERR_DEVICE;
        if (rc == ACC_OK && hipEventRecord(sl.ev, sl.st) != hipSuccess) rc = ACC_ERR_DEVICE;
        if (rc != ACC_OK) {  // fail these requests; the slot stays free
            const auto now = Clock::now();
            for (uint64_t id : sl.ids) {
                auto it = s->reqs.find(id);
                if (it == s->reqs.end()) continue;
                Request &r = it->second;
                r.done = true;
                r.status = rc;
                r.t1 = now;
                ++s->completed;
                if (!r.keep) erase_request(s, it);  // nobody waits for a load-generator request
            }
            sl.ids.clear();
            s->cv_done.notify_all();
            continue;
        }
        ++mx.inflight;
        s->peak_inflight = std::max<int64_t>(s->peak_inflight, mx.inflight);
        sl.busy = true;
        ++s->batches;
        s->batched += nq;
    }
}

void worker_main(acc_server *s) {
    hipSetDevice(s->o.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        const int64_t before = s->completed;
        reap(s);
        if (s->completed != before) s->cv_done.notify_all();
        launch(s);
        bool busy = false;
        for (int i = 0; i < s->o.slots; ++i) busy |= s->slots[i].busy;
        if (busy)
            s->cv_work.wait_for(lk, std::chrono::microseconds(20));  // poll the in-flight batches
        else
            s->cv_work.wait(lk, [s] { return s->stop || s->npending > 0; });
    }
}

void free_slots(acc_server *s) {
    for (Slot &sl : s->slots) {
        acc::batch_ws_destroy(sl.ws, s->o.device);
        if (sl.ev) hipEventDestroy(sl.ev);
        if (sl.st) hipStreamDestroy(sl.st);
        if (sl.x_pin) hipHostFree(sl.x_pin);
        if (sl.o_pin) hipHostFree(sl.o_pin);
        if (sl.x_dev) hipFree(sl.x_dev);
        if (sl.o_dev) hipFree(sl.o_dev);
        sl = Slot();
    }
}

}  // namespace

extern "C" int acc_server_create(const acc_server_opts *opts, acc_server **out) {
    if (!opts || !out) return fail(ACC_ERR_ARG, "null argument");
    acc_server_opts o = *opts;
    if (o.max_batch == 0) o.max_batch = 256;
    if (o.slots == 0) o.slots = 4;
    if (o.ck == 0) o.ck = 512;
    if (o.max_pending == 0) o.max_pending = 1 << 16;
    if (o.max_request_bytes == 0) o.max_request_bytes = int64_t(1) << 30;
    if (o.max_batch < 1 || o.max_batch > 1024 || o.slots < 1 || o.slots > kMaxSlots || o.ck < 1 || o.max_pending < 1 ||
        o.memory_budget < 0 || o.max_request_bytes < 0)
        return fail(ACC_ERR_ARG,
                    "max_batch in [1, 1024], slots in [1, 8], ck >= 1, max_pending >= 1, budgets >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(ACC_ERR_NODEVICE, "no GPU visible");
    if (o.device < 0 || o.device >= ndev) return fail(ACC_ERR_ARG, "bad device index");
    if (hipSetDevice(o.device) != hipSuccess) return fail(ACC_ERR_DEVICE, "hipSetDevice failed");
    if (o.memory_budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(ACC_ERR_DEVICE, "hipMemGetInfo failed");
        o.memory_budget = (int64_t)(fr / 10 * 9);
    }
    acc_server *s = new acc_server();
    s->o = o;
    s->mem.set_limit(o.memory_budget);
    int rc = ACC_OK;
    const size_t xb = (size_t)o.max_batch * s->max_cols * 4, ob = (size_t)o.max_batch * kMaxK * sizeof(acc_element);
    if (!s->mem.reserve((int64_t)(xb + ob) * o.slots)) rc = fail(ACC_ERR_BUDGET, "memory budget below the slot buffers");
    for (int i = 0; i < o.slots && rc == ACC_OK; ++i) {
        Slot &sl = s->slots[i];
        sl.ws = acc::batch_ws_create();
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&sl.x_pin, xb, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&sl.o_pin, ob, hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void **)&sl.x_dev, xb) != hipSuccess || hipMalloc((void **)&sl.o_dev, ob) != hipSuccess)
            rc = fail(ACC_ERR_NOMEM, "server slot allocation failed");
    }
    if (rc != ACC_OK) {
        free_slots(s);
        delete s;
        return rc;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return ACC_OK;
}

// Device bytes acc_gpu_upload will hold for this matrix, before the
// workspaces: the FP32 CSR (4-byte row pointers), the packets with their
// padding, metadata and slot map, and per-row scores.
static int64_t upload_estimate(const acc_csr *csr, const acc_packed *pk) {
    const int64_t nnz = csr->nnz();
    int64_t b = ((int64_t)csr->rows + 1) * 4 + nnz * 8 + (int64_t)csr->rows * 4;
    b += (int64_t)pk->words.size() * 4 + (int64_t)pk->nslices() * (8 + 4 + 1 + 256) + 32 * 64 * 4;
    return b;
}

extern "C" int acc_server_add_matrix(acc_server *s, const acc_csr *csr, const acc_packed *pk, int *matrix_id) {
    if (!s || !csr || !pk || !matrix_id) return fail(ACC_ERR_ARG, "a server, the FP32 CSR and its packets are required");
    if (csr->cols > s->max_cols) return fail(ACC_ERR_ARG, "cols must be <= 8192");
    if (hipSetDevice(s->o.device) != hipSuccess) return fail(ACC_ERR_DEVICE, "hipSetDevice failed");
    // admission: the matrix's bytes are reserved under the lock before the
    // upload, so concurrent additions cannot both pass against the same room;
    // the reservation is corrected to the measured bytes afterwards
    const int64_t reserved = upload_estimate(csr, pk);
    if (!s->mem.reserve(reserved)) return fail(ACC_ERR_BUDGET, "matrix does not fit in the server's memory budget");
    auto release = [&](int rc) {
        s->mem.release(reserved);
        return rc;
    };
    // the room this matrix may grow into: what is free plus its reservation
    const int64_t have = s->mem.limit() - s->mem.used() + reserved;
    acc_gpu_matrix *m = nullptr;
    int rc = acc_gpu_upload(csr, pk, s->o.device, &m);
    if (rc != ACC_OK) return release(rc);
    // the resident matrix plus the workspace of a full batch and of a small one
    acc_gpu_info_t info;
    acc_gpu_matrix_info(m, &info);
    if (info.device_bytes > have) {
        acc_gpu_matrix_destroy(m);
        return release(fail(ACC_ERR_BUDGET, "matrix does not fit in the server's memory budget"));
    }
    if (rc == ACC_OK && s->o.autotune)
        rc = acc_gpu_autotune(m, s->o.ck, std::min(100, std::min(kMaxK, s->o.ck)), s->o.max_batch, nullptr, 0, nullptr);
    // the workspace a slot needs for this matrix: a scratch workspace warmed
    // with batches of 1, 2, 4, ... max_batch queries (every layout the
    // handle's batch-size table can pick), then freed
    int64_t need = 0;
    {
        acc::BatchWorkspace *w = acc::batch_ws_create();
        float *x = nullptr;
        acc_element *o = nullptr;
        if (rc == ACC_OK && (hipMalloc((void **)&x, (size_t)s->o.max_batch * csr->cols * 4) != hipSuccess ||
                             hipMalloc((void **)&o, (size_t)s->o.max_batch * sizeof(acc_element)) != hipSuccess))
            rc = fail(ACC_ERR_NOMEM, "warm-up buffers");
        if (rc == ACC_OK && hipMemset(x, 0, (size_t)s->o.max_batch * csr->cols * 4) != hipSuccess) rc = ACC_ERR_DEVICE;
        std::vector<int32_t> ks((size_t)s->o.max_batch, 1);
        for (int b = 1; rc == ACC_OK; b = std::min(2 * b, s->o.max_batch)) {
            rc = acc::topk_lowbit_batch_ws(m, w, x, b, ks.data(), s->o.ck, o, 1, nullptr);
            if (rc == ACC_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(ACC_ERR_DEVICE, "warm-up batch failed");
            if (b == s->o.max_batch) break;
        }
        need = acc::batch_ws_bytes(w);
        acc::batch_ws_destroy(w, s->o.device);
        if (x) hipFree(x);
        if (o) hipFree(o);
    }
    if (rc == ACC_OK) acc_gpu_matrix_info(m, &info);
    if (rc != ACC_OK) {
        acc_gpu_matrix_destroy(m);
        return release(rc);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    // the reservation becomes the measured bytes plus every slot's growth to
    // this matrix's workspace, if they still fit (under the lock, so that
    // concurrent additions see each other's workspace growth)
    const int64_t grow = std::max<int64_t>(0, need - s->ws_need) * s->o.slots;
    if (!s->mem.resize(reserved, info.device_bytes + grow)) {
        acc_gpu_matrix_destroy(m);
        s->mem.release(reserved);
        return fail(ACC_ERR_BUDGET, "matrix and the slots' batch workspaces exceed the budget");
    }
    s->ws_need = std::max(s->ws_need, need);
    Matrix mx;
    mx.m = m;
    mx.cols = csr->cols;
    mx.bytes = info.device_bytes + grow;
    s->mats.push_back(std::move(mx));
    *matrix_id = (int)s->mats.size() - 1;
    return ACC_OK;
}

static int submit(acc_server *s, int matrix_id, const float *x, int k, bool keep, Clock::time_point t0,
                  uint64_t *ticket) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (matrix_id < 0 || matrix_id >= (int)s->mats.size() || !s->mats[(size_t)matrix_id].m)
        return fail(ACC_ERR_ARG, "unknown matrix");
    if (k < 1 || k > std::min(kMaxK, s->o.ck)) return fail(ACC_ERR_ARG, "k must be in [1, min(960, ck)]");
    if (s->npending >= s->o.max_pending) {
        ++s->refused;
        return fail(ACC_ERR_BUDGET, "request queue full");
    }
    Matrix &mx = s->mats[(size_t)matrix_id];
    // host memory the request holds until it is read: the query, the
    // results and its bookkeeping
    const int64_t bytes = (int64_t)sizeof(Request) + 64 + (int64_t)mx.cols * 4 + (int64_t)k * (int64_t)sizeof(acc_element);
    if (s->req_bytes + bytes > s->o.max_request_bytes) {
        ++s->refused;
        return fail(ACC_ERR_BUDGET, "request memory budget full");
    }
    const uint64_t id = s->next_id++;
    Request &r = s->reqs[id];
    r.bytes = bytes;
    s->req_bytes += bytes;
    r.matrix = matrix_id;
    r.k = k;
    r.keep = keep;
    r.x.assign(x, x + mx.cols);
    r.t0 = t0;
    mx.pending.push_back(id);
    ++s->npending;
    ++s->submitted;
    if (ticket) *ticket = id;
    s->cv_work.notify_one();
    return ACC_OK;
}

extern "C" int acc_server_submit(acc_server *s, int matrix_id, const float *x, int k, uint64_t *ticket) {
    if (!s || !x || !ticket) return fail(ACC_ERR_ARG, "null argument");
    return submit(s, matrix_id, x, k, true, Clock::now(), ticket);
}

extern "C" int acc_server_wait(acc_server *s, uint64_t ticket, acc_element *out, int timeout_ms) {
    if (!s) return fail(ACC_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(ACC_ERR_ARG, "unknown ticket");
    auto ready = [&] { return s->reqs[ticket].done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return fail(ACC_ERR_BUDGET, "timed out");
    it = s->reqs.find(ticket);
    Request &r = it->second;
    const int status = r.status;
    if (status == ACC_OK && out) std::copy(r.out.begin(), r.out.end(), out);
    erase_request(s, it);
    return status == ACC_OK ? ACC_OK : fail(status, "request failed on the device");
}

static void fill_stats(acc_server *s, acc_server_stats_t *st) {
    std::memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->refused = s->refused;
    st->batches = s->batches;
    st->mean_batch = s->batches ? (double)s->batched / (double)s->batches : 0.0;
    std::vector<double> l = s->lat_us;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        auto pct = [&](double q) { return l[std::min(l.size() - 1, (size_t)(q * (double)(l.size() - 1) + 0.5))]; };
        st->p50_us = pct(0.50);
        st->p99_us = pct(0.99);
        st->max_us = l.back();
    }
    const double secs = std::chrono::duration<double>(Clock::now() - s->t_reset).count();
    st->qps = secs > 0 ? (double)s->completed / secs : 0.0;
    st->device_bytes = s->mem.used();
    st->request_bytes = s->req_bytes;
    st->peak_inflight = s->peak_inflight;
}

extern "C" int acc_server_stats(acc_server *s, acc_server_stats_t *st) {
    if (!s || !st) return fail(ACC_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    fill_stats(s, st);
    return ACC_OK;
}

extern "C" int acc_server_reset_stats(acc_server *s) {
    if (!s) return fail(ACC_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_us.clear();
    s->submitted = s->completed = s->refused = s->batches = s->batched = 0;
    s->peak_inflight = 0;
    s->t_reset = Clock::now();
    return ACC_OK;
}

extern "C" int acc_server_bench_poisson(acc_server *s, double rate_qps, double seconds, int kmax, uint64_t seed,
                                        acc_server_stats_t *st) {
    if (!s || !st || rate_qps <= 0 || seconds <= 0 || kmax < 1) return fail(ACC_ERR_ARG, "bad load parameters");
    int nmat = 0, cols_max = 0;
    std::vector<int> cols;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        nmat = (int)s->mats.size();
        for (auto &m : s->mats) cols.push_back(m.cols);
    }
    if (nmat == 0) return fail(ACC_ERR_ARG, "no matrices");
    for (int c : cols) cols_max = std::max(cols_max, c);
    // a pool of N(0,1) queries per matrix; arrivals: exponential gaps
    constexpr int kPool = 256;
    std::vector<std::vector<float>> pool((size_t)nmat);
    for (int i = 0; i < nmat; ++i) {
        pool[(size_t)i].resize((size_t)kPool * cols[(size_t)i]);
        acc_generate_queries(kPool, cols[(size_t)i], ACC_VAL_NORMAL, 0, seed + (uint64_t)i, pool[(size_t)i].data());
    }
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate_qps);
    std::uniform_int_distribution<int> pick_m(0, nmat - 1), pick_q(0, kPool - 1), pick_k(1, std::min(kmax, std::min(kMaxK, s->o.ck)));
    acc_server_reset_stats(s);
    const Clock::time_point start = Clock::now();
    double t = 0.0;
    int64_t refused = 0;
    int rc = ACC_OK;
    while (t < seconds && rc == ACC_OK) {
        t += gap(rng);
        const Clock::time_point due = start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t));
        while (Clock::now() < due) {
        }
        const int mi = pick_m(rng);
        const float *x = pool[(size_t)mi].data() + (size_t)pick_q(rng) * cols[(size_t)mi];
        uint64_t id = 0;
        // latency counts from the scheduled arrival (no coordinated omission)
        const int r = submit(s, mi, x, pick_k(rng), false, due, &id);
        if (r == ACC_ERR_BUDGET) ++refused;
        else if (r != ACC_OK) rc = r;
    }
    // drain
    for (;;) {
        std::unique_lock<std::mutex> lk(s->mu);
        bool idle = s->npending == 0;
        for (int i = 0; i < s->o.slots; ++i) idle = idle && !s->slots[i].busy;
        if (idle) {
            fill_stats(s, st);
            const double secs = std::chrono::duration<double>(Clock::now() - start).count();
            st->qps = secs > 0 ? (double)s->completed / secs : 0.0;
            break;
        }
        s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
    }
    (void)refused;
    return rc;
}

extern "C" void acc_server_destroy(acc_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    hipSetDevice(s->o.device);
    hipDeviceSynchronize();
    for (auto &m : s->mats)
        if (m.m) acc_gpu_matrix_destroy(m.m);
    free_slots(s);
    delete s;
}
