// Continuous-batching reconstruction server over resident matrices (the
// serving pattern of many MLEM requests against a few system matrices; the
// reference has one reconstruction per process, main.cu:42-143).
//
// Each registered matrix gets a batch of slots (mlem_batch_*).  Requests are
// queued per matrix; a step admits queued requests into free slots (a
// request joins its batch between two iterations), runs one MLEM iteration
// of every batch that has active slots and none in flight (each batch on its
// own stream, so batches of different matrices overlap on the device and go
// at their own pace), retires the requests that reached their iteration
// count (image copied out, slot freed) and records each request's latency
// from submission.  A request's image does not depend on what else is in
// its batch.
//
// HBM budget: a matrix is admitted with its tiles and its batch's fixed
// buffers; a request is admitted into a slot of a pass that holds memory
// already, or, when that pass's vectors (BC slots' counts, ratios and
// images) fit in what is left of the budget, into a new one; otherwise it
// waits, and passes of other matrices that have no active slot are freed to
// make room.  Admission and retirement queue their copies on the batch's
// stream without waiting (one wait per step that retires requests, for the
// images); a request's inputs are kept until its first iteration has ended.
// Requests move to lower passes when that empties a pass (compact below).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "abi_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

int fail(int st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    mlem::set_last_error(buf);
    return st;
}

struct Request {
    uint64_t id = 0;
    uint32_t matrix = 0, iters = 0, done_iters = 0;
    std::vector<int32_t> g;  // inputs, kept until the first iteration has ended (their copies are queued)
    std::vector<float> f0, image;
    Clock::time_point submitted, finished;
    bool finished_flag = false;
};

struct Served {
    mlem_matrix *m = nullptr;
    mlem_batch *batch = nullptr;
    uint32_t rows = 0, cols = 0, slots = 0;
    uint64_t bytes = 0;                     // matrix + batch, fixed part
    std::vector<uint64_t> slot_req;         // request id per slot, 0 = free
    std::deque<uint64_t> queue;             // waiting requests
    bool inflight = false;                  // an iteration is running on the batch's stream
};

}  // namespace

struct mlem_server {
    uint64_t budget = 0, used = 0;
    mlem_batch_opts bopts{};
    std::vector<std::unique_ptr<Served>> mats;
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    uint64_t steps = 0;
    uint64_t pass_sweeps = 0, slot_moves = 0;
};

extern "C" {

int mlem_server_create(uint64_t hbm_budget_bytes, const mlem_batch_opts *batch_opts, mlem_server **out) {
    if (!out) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_create: null out");
    *out = nullptr;
    auto s = std::make_unique<mlem_server>();
    s->budget = hbm_budget_bytes;
    if (batch_opts)
        s->bopts = *batch_opts;
    else
        mlem_batch_opts_default(&s->bopts);
    *out = s.release();
    return MLEM_OK;
}

int mlem_server_add_matrix(mlem_server *s, mlem_matrix *m, uint32_t max_slots, uint32_t *matrix_id) {
    if (!s || !m || !matrix_id) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_add_matrix: null argument");
    uint32_t rows = 0, cols = 0;
    uint64_t nnz = 0, mbytes = 0;
    int has_t = 0;
    if (int st = mlem_matrix_info(m, &rows, &cols, &nnz, &has_t, &mbytes))
        return fail(st, "mlem_server_add_matrix: %s", std::string(mlem_last_error()).c_str());
    if (!has_t && !m->tiles_only)
        return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_add_matrix: the matrix needs its transpose");
    if (max_slots == 0) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_add_matrix: no slots");
    // The batch (its tiles included) is built first; it is kept only when
    // the matrix, the batch's fixed buffers and one pass fit in what is left
    // of the budget.
    mlem_batch *b = nullptr;
    if (int st = mlem_batch_create(m, max_slots, &s->bopts, &b))
        return fail(st, "mlem_server_add_matrix: %s", std::string(mlem_last_error()).c_str());
    const uint64_t bbytes = b->fixed_bytes(), pbytes = b->pass_bytes();
    mlem_matrix_info(m, nullptr, nullptr, nullptr, nullptr, &mbytes);  // with the tiles the batch built
    if (s->budget && s->used + mbytes + bbytes + pbytes > s->budget) {
        mlem_batch_destroy(b);
        return fail(MLEM_ERR_OUT_OF_MEMORY,
                    "mlem_server_add_matrix: %.3f GB (matrix %.3f + batch %.3f + one pass %.3f) exceed the %.3f GB "
                    "left of the budget",
                    (mbytes + bbytes + pbytes) / 1e9, mbytes / 1e9, bbytes / 1e9, pbytes / 1e9,
                    (s->budget - s->used) / 1e9);
    }
    auto sv = std::make_unique<Served>();
    sv->m = m;
    sv->batch = b;
    sv->rows = rows;
    sv->cols = cols;
    sv->slots = max_slots;
    sv->bytes = mbytes + bbytes;
    sv->slot_req.assign(max_slots, 0);
    s->used += sv->bytes;
    *matrix_id = (uint32_t)s->mats.size();
    s->mats.push_back(std::move(sv));
    return MLEM_OK;
}

int mlem_server_submit(mlem_server *s, uint32_t matrix_id, const int32_t *g, const float *f0, uint32_t iters,
                       uint64_t *request_id) {
    if (!s || !g || !f0 || !request_id) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_submit: null argument");
    if (matrix_id >= s->mats.size()) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_submit: no matrix %u", matrix_id);
    if (iters == 0) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_submit: iters must be > 0");
    Served &sv = *s->mats[matrix_id];
    auto r = std::make_unique<Request>();
    r->id = s->next_id++;
    r->matrix = matrix_id;
    r->iters = iters;
    r->g.assign(g, g + sv.rows);
    r->f0.assign(f0, f0 + sv.cols);
    r->submitted = Clock::now();
    sv.queue.push_back(r->id);
    *request_id = r->id;
    s->reqs.emplace(r->id, std::move(r));
    return MLEM_OK;
}

}  // extern "C"

namespace {

// Frees passes without active slots of the batches other than `keep` that
// have no iteration in flight, until `need` bytes fit in the budget; true
// when they do.
bool make_room(mlem_server *s, const Served *keep, uint64_t need) {
    for (auto &svp : s->mats) {
        if (s->used + need <= s->budget) break;
        Served &o = *svp;
        if (&o == keep || o.inflight) continue;
        for (uint32_t p = 0; p < o.batch->passes && s->used + need > s->budget; ++p)
            s->used -= mlem::abi::batch_release_pass(o.batch, p);
    }
    return s->used + need <= s->budget;
}

// Queued requests into free slots: first slots of passes that hold memory,
// then slots of new passes while they fit in the budget.
int admit(mlem_server *s, Served &sv) {
    mlem_batch *b = sv.batch;
    while (!sv.queue.empty()) {
        int64_t pick = -1;
        for (uint32_t slot = 0; slot < sv.slots && pick < 0; ++slot)
            if (!sv.slot_req[slot] && b->pass[slot / b->BC]) pick = slot;
        if (pick < 0) {
            for (uint32_t slot = 0; slot < sv.slots && pick < 0; ++slot)
                if (!sv.slot_req[slot]) pick = slot;
            if (pick < 0) break;  // every slot busy
            if (s->budget && s->used + b->pass_bytes() > s->budget && !make_room(s, &sv, b->pass_bytes()))
                break;  // wait for memory
        }
        Request &r = *s->reqs.at(sv.queue.front());
        sv.queue.pop_front();
        uint64_t allocated = 0;
        if (int st = mlem::abi::batch_set_async(b, (uint32_t)pick, r.g.data(), r.f0.data(), &allocated))
            return fail(st, "mlem_server_step: %s", std::string(mlem_last_error()).c_str());
        s->used += allocated;
        sv.slot_req[pick] = r.id;
    }
    return MLEM_OK;
}

// Every pass with an active slot costs a full sweep of the matrix, however
// few of its slots are active: while the requests of the highest pass that
// has any fit in the free slots of lower passes that hold memory, they move
// there (lowest slot first), so the batch sweeps as few passes as its
// requests need (after retirements the survivors are scattered; in a
// backlog of 16 requests per matrix with 10-100 iterations each, about a
// quarter of the sweeps go otherwise).
int compact(mlem_server *s, Served &sv) {
    mlem_batch *b = sv.batch;
    const uint32_t BC = b->BC;
    for (;;) {
        int64_t hi = -1;
        for (uint32_t slot = 0; slot < sv.slots; ++slot)
            if (sv.slot_req[slot]) hi = slot / BC;
        if (hi <= 0) return MLEM_OK;
        std::vector<uint32_t> from, to;
        for (uint32_t slot = (uint32_t)hi * BC; slot < std::min(sv.slots, (uint32_t)(hi + 1) * BC); ++slot)
            if (sv.slot_req[slot]) from.push_back(slot);
        for (uint32_t slot = 0; slot < (uint32_t)hi * BC && to.size() < from.size(); ++slot)
            if (!sv.slot_req[slot] && b->pass[slot / BC]) to.push_back(slot);
        if (to.size() < from.size()) return MLEM_OK;  // the pass cannot be emptied
        for (size_t k = 0; k < from.size(); ++k) {
            if (int st = mlem::abi::batch_move_slot_async(b, from[k], to[k]))
                return fail(st, "mlem_server_step: %s", std::string(mlem_last_error()).c_str());
            sv.slot_req[to[k]] = sv.slot_req[from[k]];
            sv.slot_req[from[k]] = 0;
            ++s->slot_moves;
        }
    }
}

// The iteration in flight on sv's stream has ended: count it for every
// request in the batch and retire those that reached their count (their
// images copied out with one wait for all of them).
int complete(mlem_server *s, Served &sv) {
    sv.inflight = false;
    std::vector<Request *> done;
    for (uint32_t slot = 0; slot < sv.slots; ++slot) {
        const uint64_t id = sv.slot_req[slot];
        if (!id) continue;
        Request &r = *s->reqs.at(id);
        if (r.done_iters == 0) {  // its input copies ran before this iteration
            std::vector<int32_t>().swap(r.g);
            std::vector<float>().swap(r.f0);
        }
        if (++r.done_iters < r.iters) continue;
        r.image.resize(sv.cols);
        if (int st = mlem::abi::batch_get_image_async(sv.batch, slot, r.image.data()))
            return fail(st, "mlem_server_step: %s", std::string(mlem_last_error()).c_str());
        if (int st = mlem::abi::batch_deactivate_async(sv.batch, slot))
            return fail(st, "mlem_server_step: %s", std::string(mlem_last_error()).c_str());
        sv.slot_req[slot] = 0;
        done.push_back(&r);
    }
    if (!done.empty()) {
        const hipError_t e = hipStreamSynchronize(sv.batch->stream);
        if (e != hipSuccess) return fail(MLEM_ERR_HIP, "mlem_server_step: %s", hipGetErrorString(e));
        const Clock::time_point now = Clock::now();
        for (Request *r : done) {
            r->finished = now;
            r->finished_flag = true;
        }
    }
    return MLEM_OK;
}

}  // namespace

extern "C" {

// One scheduling step.  Every batch with no iteration in flight admits queued
// requests into its free slots and, if it has active slots, starts one
// iteration on its stream; then the step waits up to max_wait_us for an
// iteration to end (0: only collects those already ended) and retires the
// requests that reached their count.  Batches of different matrices thus run
// at their own pace: a small matrix iterates many times while a large one is
// in flight.  *active = requests in slots after the step, *queued = waiting.
int mlem_server_step(mlem_server *s, uint32_t max_wait_us, uint32_t *active, uint32_t *queued) {
    if (!s) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_step: null server");
    bool any_inflight = false;
    for (auto &svp : s->mats) {
        Served &sv = *svp;
        if (!sv.inflight) {
            if (int st = admit(s, sv)) return st;
            if (int st = compact(s, sv)) return st;
            bool any = false;
            for (uint32_t p = 0, BC = sv.batch->BC; p * BC < sv.slots; ++p) {
                bool used = false;
                for (uint32_t slot = p * BC; slot < std::min(sv.slots, (p + 1) * BC); ++slot)
                    used = used || sv.slot_req[slot] != 0;
                s->pass_sweeps += used ? 1u : 0u;
                any = any || used;
            }
            if (any) {
                if (int st = mlem_batch_iterate(sv.batch, 1))
                    return fail(st, "mlem_server_step: %s", std::string(mlem_last_error()).c_str());
                sv.inflight = true;
            }
        }
        any_inflight = any_inflight || sv.inflight;
    }
    ++s->steps;
    const auto deadline = Clock::now() + std::chrono::microseconds(max_wait_us);
    for (bool ended = false; any_inflight && !ended;) {
        for (auto &svp : s->mats) {
            Served &sv = *svp;
            if (!sv.inflight) continue;
            const hipError_t q = hipStreamQuery(sv.batch->stream);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return fail(MLEM_ERR_HIP, "mlem_server_step: %s", hipGetErrorString(q));
            if (int st = complete(s, sv)) return st;
            ended = true;
        }
        if (ended || Clock::now() >= deadline) break;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    uint32_t act = 0, que = 0;
    for (auto &svp : s->mats) {
        for (uint64_t id : svp->slot_req) act += id != 0;
        que += (uint32_t)svp->queue.size();
    }
    if (active) *active = act;
    if (queued) *queued = que;
    return MLEM_OK;
}

// 1 and the image (f32[cols]) and latency when the request is done (the
// request is then forgotten), 0 while it is queued or running.
int mlem_server_poll(mlem_server *s, uint64_t request_id, float *f_out, double *latency_s, int *done) {
    if (!s || !done) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_poll: null argument");
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_poll: unknown request %llu",
                                         (unsigned long long)request_id);
    Request &r = *it->second;
    *done = r.finished_flag ? 1 : 0;
    if (!r.finished_flag) return MLEM_OK;
    if (f_out) std::memcpy(f_out, r.image.data(), sizeof(float) * r.image.size());
    if (latency_s) *latency_s = std::chrono::duration<double>(r.finished - r.submitted).count();
    s->reqs.erase(it);
    return MLEM_OK;
}

int mlem_server_info(const mlem_server *s, uint64_t *budget, uint64_t *used, uint32_t *matrices, uint64_t *steps) {
    if (!s) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_info: null server");
    if (budget) *budget = s->budget;
    if (used) *used = s->used;
    if (matrices) *matrices = (uint32_t)s->mats.size();
    if (steps) *steps = s->steps;
    return MLEM_OK;
}

int mlem_server_stats(const mlem_server *s, uint64_t *pass_sweeps, uint64_t *slot_moves) {
    if (!s) return fail(MLEM_ERR_INVALID_ARGUMENT, "mlem_server_stats: null server");
    if (pass_sweeps) *pass_sweeps = s->pass_sweeps;
    if (slot_moves) *slot_moves = s->slot_moves;
    return MLEM_OK;
}

int mlem_server_destroy(mlem_server *s) {
    if (!s) return MLEM_OK;  // (mlem_batch_destroy waits for the batch's stream)
    for (auto &sv : s->mats) mlem_batch_destroy(sv->batch);
    delete s;
    return MLEM_OK;
}

}  // extern "C"
