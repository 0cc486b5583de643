// Asynchronous writer of the gk2.istate* / gq2.istate* files
// (src/interactions.f90:2337-2349, 1877-1889) from device rows: submit()
// queues a device -> pinned-host copy of the rows behind the work already on
// the context's stream, makes later work on that stream wait for the copy (so
// the caller may reuse its device buffer for the next states at once) and
// returns; a host thread writes the files while the device computes.  Two
// pinned staging buffers alternate, so one batch can be copied while the
// previous one is being written.
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ctx.h"

using elph::Error;

struct elph_g2_writer {
    elph_ctx *ctx = nullptr;
    std::string dir;
    char kind = 'k';
    int64_t stride = 0, max_states = 0;
    hipStream_t stream = nullptr;
    double *buf[2] = {nullptr, nullptr};
    hipEvent_t ready[2] = {nullptr, nullptr}, copied[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};
    int next = 0;
    struct Job {
        int b;
        int64_t first;
        std::vector<int64_t> counts;
    };
    std::deque<Job> jobs;
    int in_flight = 0;
    int status = ELPH_OK;
    std::string err;
    bool stop = false;
    std::mutex mu;
    std::condition_variable cv;
    std::thread th;

    void work() {
        for (;;) {
            Job j;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !jobs.empty(); });
                if (jobs.empty()) return;
                j = std::move(jobs.front());
                jobs.pop_front();
            }
            int st = ELPH_OK;
            if (hipEventSynchronize(copied[j.b]) != hipSuccess) st = ELPH_EDEVICE;
            if (st == ELPH_OK)
                st = elph_write_g2_files(dir.c_str(), kind, j.first, (int64_t)j.counts.size(), j.counts.data(),
                                         buf[j.b], stride);
            {
                std::lock_guard<std::mutex> lk(mu);
                busy[j.b] = false;
                --in_flight;
                if (st != ELPH_OK && status == ELPH_OK) {
                    status = st;
                    err = st == ELPH_EIO ? "writing g2 files into " + dir + " failed" : "device copy failed";
                }
            }
            cv.notify_all();
        }
    }
};

namespace {
void writer_release(elph_g2_writer *w) {
    for (int b = 0; b < 2; ++b) {
        if (w->buf[b]) (void)hipHostFree(w->buf[b]);
        if (w->ready[b]) (void)hipEventDestroy(w->ready[b]);
        if (w->copied[b]) (void)hipEventDestroy(w->copied[b]);
    }
    if (w->stream) (void)hipStreamDestroy(w->stream);
}
}  // namespace

extern "C" {

int elph_g2_writer_create(elph_ctx *ctx, const char *dir, char kind, int64_t row_stride, int64_t max_states,
                          elph_g2_writer **out) {
    if (!ctx || !dir || !out || row_stride <= 0 || max_states <= 0 || (kind != 'k' && kind != 'q'))
        return ELPH_EINVAL;
    *out = nullptr;
    auto *w = new elph_g2_writer();
    w->ctx = ctx;
    w->dir = dir;
    w->kind = kind;
    w->stride = row_stride;
    w->max_states = max_states;
    try {
        HCHECK(hipSetDevice(ctx->device));
        HCHECK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
        for (int b = 0; b < 2; ++b) {
            HCHECK(hipHostMalloc((void **)&w->buf[b], (size_t)max_states * row_stride * sizeof(double),
                                 hipHostMallocDefault));
            HCHECK(hipEventCreateWithFlags(&w->ready[b], hipEventDisableTiming));
            HCHECK(hipEventCreateWithFlags(&w->copied[b], hipEventDisableTiming));
        }
    } catch (const Error &e) {
        writer_release(w);
        delete w;
        return elph::ret_error(ctx, e);
    }
    w->th = std::thread([w] { w->work(); });
    *out = w;
    return ELPH_OK;
}

int elph_g2_writer_submit(elph_g2_writer *w, int64_t istate_first, int64_t nstates, const int64_t *counts,
                          const double *g2_device) {
    if (!w || !counts || !g2_device || nstates < 0 || nstates > w->max_states) return ELPH_EINVAL;
    for (int64_t t = 0; t < nstates; ++t)
        if (counts[t] > w->stride) return ELPH_EINVAL;
    int b;
    {
        std::unique_lock<std::mutex> lk(w->mu);
        if (w->status != ELPH_OK) return w->status;
        b = w->next;
        w->cv.wait(lk, [&] { return !w->busy[b]; });  // its previous batch is on disk
        w->busy[b] = true;
        w->next = 1 - b;
        ++w->in_flight;
    }
    try {
        elph_ctx *c = w->ctx;
        HCHECK(hipSetDevice(c->device));
        // the rows are complete once the work queued so far on the context's stream is
        HCHECK(hipEventRecord(w->ready[b], c->stream));
        HCHECK(hipStreamWaitEvent(w->stream, w->ready[b], 0));
        for (int64_t t = 0; t < nstates; ++t) {
            if (counts[t] < 0) continue;
            const int64_t n = w->kind == 'k' ? w->stride : counts[t];
            if (n > 0)
                HCHECK(hipMemcpyAsync(w->buf[b] + t * w->stride, g2_device + t * w->stride, n * sizeof(double),
                                      hipMemcpyDeviceToHost, w->stream));
        }
        HCHECK(hipEventRecord(w->copied[b], w->stream));
        // later work on the context's stream (the next run into the same rows) waits for the copy
        HCHECK(hipStreamWaitEvent(c->stream, w->copied[b], 0));
    } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(w->mu);
        w->busy[b] = false;
        --w->in_flight;
        if (w->status == ELPH_OK) {
            w->status = e.code;
            w->err = e.msg;
        }
        w->cv.notify_all();
        return e.code;
    }
    {
        std::lock_guard<std::mutex> lk(w->mu);
        w->jobs.push_back({b, istate_first, std::vector<int64_t>(counts, counts + nstates)});
    }
    w->cv.notify_all();
    return ELPH_OK;
}

int elph_g2_writer_wait(elph_g2_writer *w) {
    if (!w) return ELPH_EINVAL;
    std::unique_lock<std::mutex> lk(w->mu);
    w->cv.wait(lk, [&] { return w->in_flight == 0; });
    return w->status;
}

const char *elph_g2_writer_error(const elph_g2_writer *w) { return w ? w->err.c_str() : "no writer"; }

int elph_g2_writer_destroy(elph_g2_writer *w) {
    if (!w) return ELPH_OK;
    const int st = elph_g2_writer_wait(w);
    {
        std::lock_guard<std::mutex> lk(w->mu);
        w->stop = true;
    }
    w->cv.notify_all();
    if (w->th.joinable()) w->th.join();
    (void)hipSetDevice(w->ctx->device);
    writer_release(w);
    delete w;
    return st;
}

}  // extern "C"
