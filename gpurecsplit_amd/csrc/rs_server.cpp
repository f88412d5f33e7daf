// rs_server.cpp -- construction server: asynchronous jobs of differing size and
// leaf size, built with continuous batching on one GPU.
//
// Lifecycle of a job:
//   submit   the keys are copied into the server's device key pool at once
//            (device-resident from arrival); if the HBM budget cannot hold them,
//            the job waits in host memory and is uploaded as memory frees up
//   ready    a free build slot takes the oldest ready job together with every
//            other ready job of the same (leaf size, bucket size, rotation), up to
//            max_batch_keys, and builds them as one batch (one bucket space, one
//            persistent search launch per tree level for all of them); the batch
//            starts only when its worst-case working set fits the budget
//   done     the job's MPHF is handed out by rs_wait; its keys leave the pool
// Slots are build contexts with their own HIP stream, so the batches of different
// slots (other leaf sizes, or jobs that arrived while a batch was running) overlap
// on the device.  HBM is accounted from real allocations: the key pool's chunks
// plus every slot context's buffers, plus the growth bound of running batches.
// A batch that fails is retried job by job, so only a failing job reports an
// error and every result equals the job's standalone build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rs_internal.h"

namespace rsa {
namespace {

using Clock = std::chrono::steady_clock;

// Device memory for resident keys: chunks from hipMalloc, first-fit inside them.
// Chunks are kept until the server is destroyed or trimmed.
class KeyPool {
  public:
    static constexpr uint64_t kChunk = uint64_t(256) << 20;

    uint64_t reserved() const { return reserved_; }
    uint64_t used() const { return used_; }

    // Returns nullptr when no chunk has room and a new chunk would pass `limit`
    // bytes of reservation.
    uint8_t *alloc(uint64_t bytes, uint64_t limit) {
        bytes = (bytes + 255) & ~uint64_t(255);
        for (Chunk &c : chunks_)
            if (uint8_t *p = take(c, bytes)) return p;
        const uint64_t size = std::max(kChunk, bytes);
        if (reserved_ + size > limit) return nullptr;
        void *p = nullptr;
        if (hipMalloc(&p, size) != hipSuccess) return nullptr;
        chunks_.push_back(Chunk{(uint8_t *)p, size, {{0, size}}});
        reserved_ += size;
        return take(chunks_.back(), bytes);
    }

    void free(uint8_t *p, uint64_t bytes) {
        bytes = (bytes + 255) & ~uint64_t(255);
        for (Chunk &c : chunks_) {
            if (p < c.base || p >= c.base + c.size) continue;
            uint64_t off = (uint64_t)(p - c.base);
            used_ -= bytes;
            auto it = c.free.emplace(off, bytes).first;
            auto next = std::next(it);
            if (next != c.free.end() && it->first + it->second == next->first) {
                it->second += next->second;
                c.free.erase(next);
            }
            if (it != c.free.begin()) {
                auto prev = std::prev(it);
                if (prev->first + prev->second == it->first) {
                    prev->second += it->second;
                    c.free.erase(it);
                }
            }
            return;
        }
    }

    // Returns empty chunks to the device.
    void trim() {
        for (size_t i = chunks_.size(); i-- > 0;) {
            Chunk &c = chunks_[i];
            if (c.free.size() == 1 && c.free.begin()->second == c.size) {
                (void)hipFree(c.base);
                reserved_ -= c.size;
                chunks_.erase(chunks_.begin() + i);
            }
        }
    }

    ~KeyPool() {
        for (Chunk &c : chunks_) (void)hipFree(c.base);
    }

  private:
    struct Chunk {
        uint8_t *base;
        uint64_t size;
        std::map<uint64_t, uint64_t> free;  // offset -> length
    };
    uint8_t *take(Chunk &c, uint64_t bytes) {
        for (auto it = c.free.begin(); it != c.free.end(); ++it) {
            if (it->second < bytes) continue;
            const uint64_t off = it->first, len = it->second;
            c.free.erase(it);
            if (len > bytes) c.free.emplace(off + bytes, len - bytes);
            used_ += bytes;
            return c.base + off;
        }
        return nullptr;
    }
    std::vector<Chunk> chunks_;
    uint64_t reserved_ = 0, used_ = 0;
};

struct Job {
    uint64_t id = 0;
    uint32_t leaf = 0;
    uint64_t bucket = 0;
    int rotate = 1;  // flags
    uint64_t n = 0;
    std::vector<uint64_t> host_keys;  // only while waiting for key memory
    const uint64_t *d_keys = nullptr;
    uint8_t *pool_mem = nullptr;  // the pool copy (nullptr: caller-owned device keys)
    uint64_t pool_bytes = 0;
    int state = RS_JOB_QUEUED;
    int status = RS_OK;
    std::string error;
    double t_submit = 0, t_start = -1, t_done = -1;
    uint64_t batch_id = 0;
    uint32_t batch_jobs = 0;
    rs_mphf *result = nullptr;
};
using JobPtr = std::shared_ptr<Job>;

}  // namespace
}  // namespace rsa

using namespace rsa;

struct rs_server {
    rs_server_config cfg{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv;
    std::map<uint64_t, JobPtr> jobs;  // every job not yet handed out
    std::deque<JobPtr> host_wait;     // keys still in host memory
    std::deque<JobPtr> ready;         // keys resident, waiting for a slot
    KeyPool pool;
    hipStream_t upload = nullptr;
    struct Slot {
        rs_ctx *ctx = nullptr;
        uint64_t ctx_bytes = 0;  // device bytes the slot's context holds
        uint64_t growth = 0;     // bound of further growth of the running batch
        bool busy = false;
        std::thread th;
    };
    std::vector<std::unique_ptr<Slot>> slots;
    std::thread uploader;
    bool closing = false;
    uint64_t next_id = 1, next_batch = 1;
    rs_serve_stats st{};

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    uint64_t committed() const {
        uint64_t t = pool.reserved();
        for (const auto &s : slots) t += s->ctx_bytes + s->growth;
        return t;
    }
    void note_peak() { st.peak_committed_bytes = std::max(st.peak_committed_bytes, committed()); }

    // Called with mu held: reserve pool memory for a job's keys and copy them in.
    bool place_keys(const JobPtr &j, const uint64_t *src, hipMemcpyKind kind) {
        const uint64_t bytes = j->n * 16;
        const uint64_t others = committed() - pool.reserved();
        const uint64_t limit = cfg.hbm_budget_bytes > others ? cfg.hbm_budget_bytes - others : 0;
        uint8_t *p = pool.alloc(bytes, limit);
        if (!p) return false;
        if (hipMemcpyAsync(p, src, bytes, kind, upload) != hipSuccess || hipStreamSynchronize(upload) != hipSuccess) {
            pool.free(p, bytes);
            return false;
        }
        j->pool_mem = p;
        j->pool_bytes = bytes;
        j->d_keys = (const uint64_t *)p;
        note_peak();
        return true;
    }

    void finish(const JobPtr &j, int status, const std::string &err, rs_mphf *r, double t) {
        j->status = status;
        j->error = err;
        j->result = r;
        j->state = status == RS_OK ? RS_JOB_DONE : RS_JOB_FAILED;
        j->t_done = t;
        if (j->pool_mem) pool.free(j->pool_mem, j->pool_bytes);
        j->pool_mem = nullptr;
        j->d_keys = nullptr;
        if (status == RS_OK) {
            st.jobs_done++;
            st.keys_done += j->n;
        } else {
            st.jobs_failed++;
        }
    }

    // Called with mu held: the next batch a slot may start, or empty.
    std::vector<JobPtr> take_batch(Slot &slot, uint64_t &bound) {
        if (ready.empty()) return {};
        const JobPtr head = ready.front();
        std::vector<JobPtr> group;
        uint64_t keys = 0, nb = 0;
        for (const JobPtr &j : ready) {
            if (j->leaf != head->leaf || j->bucket != head->bucket || j->rotate != head->rotate) continue;
            if (!group.empty() && keys + j->n > cfg.max_batch_keys) break;
            group.push_back(j);
            keys += j->n;
            nb += std::max<uint64_t>(1, (j->n + j->bucket - 1) / j->bucket);
        }
        bound = build_bound_bytes(keys, nb);
        const uint64_t growth = bound > slot.ctx_bytes ? bound - slot.ctx_bytes : 0;
        if (committed() + growth > cfg.hbm_budget_bytes) {
            // idle slots give back their buffers first
            for (const auto &s : slots)
                if (!s->busy && s.get() != &slot && s->ctx_bytes) {
                    (void)rs_ctx_release(s->ctx);
                    s->ctx_bytes = rs_ctx_device_bytes(s->ctx);
                }
        }
        if (committed() + growth > cfg.hbm_budget_bytes) {
            bool others_busy = false;
            for (const auto &s : slots) others_busy |= s->busy && s.get() != &slot;
            if (others_busy) return {};  // wait for memory to free up
            // nothing else runs: start anyway if the job alone is what does not fit
        }
        for (const JobPtr &j : group) ready.erase(std::find(ready.begin(), ready.end(), j));
        slot.growth = growth;
        note_peak();
        return group;
    }

    void slot_loop(Slot *slot) {
        (void)hipSetDevice(cfg.device);
        for (;;) {
            std::vector<JobPtr> batch;
            uint64_t bound = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    batch = take_batch(*slot, bound);
                    if (!batch.empty() || (closing && ready.empty() && host_wait.empty())) break;
                    cv.wait_for(lk, std::chrono::milliseconds(20));
                }
                if (batch.empty()) return;
                slot->busy = true;
                const double t = now();
                const uint64_t bid = next_batch++;
                for (const JobPtr &j : batch) {
                    j->state = RS_JOB_RUNNING;
                    j->t_start = t;
                    j->batch_id = bid;
                    j->batch_jobs = (uint32_t)batch.size();
                }
                st.batches++;
            }
            std::vector<const uint64_t *> ptrs;
            std::vector<uint64_t> ns;
            for (const JobPtr &j : batch) {
                ptrs.push_back(j->d_keys);
                ns.push_back(j->n);
            }
            std::vector<rs_mphf *> out;
            int rc = build_sets_status(slot->ctx, ptrs, ns, batch[0]->leaf, batch[0]->bucket, batch[0]->rotate, out);
            std::vector<int> codes(batch.size(), rc);
            std::vector<std::string> errs(batch.size(), rc ? rs_last_error() : "");
            out.resize(batch.size(), nullptr);
            if (rc != RS_OK && batch.size() > 1) {
                // isolate the failing jobs: every job alone
                for (size_t i = 0; i < batch.size(); i++) {
                    std::vector<rs_mphf *> one;
                    codes[i] = build_sets_status(slot->ctx, {ptrs[i]}, {ns[i]}, batch[i]->leaf, batch[i]->bucket,
                                                 batch[i]->rotate, one);
                    errs[i] = codes[i] ? rs_last_error() : "";
                    out[i] = codes[i] == RS_OK ? one[0] : nullptr;
                }
                std::lock_guard<std::mutex> g(mu);
                st.batches_split++;
            }
            const uint64_t ctx_bytes = rs_ctx_device_bytes(slot->ctx);
            {
                std::lock_guard<std::mutex> g(mu);
                const double t = now();
                for (size_t i = 0; i < batch.size(); i++) finish(batch[i], codes[i], errs[i], out[i], t);
                slot->ctx_bytes = ctx_bytes;
                slot->growth = 0;
                slot->busy = false;
                note_peak();
            }
            cv.notify_all();
        }
    }

    // Moves jobs whose keys wait in host memory into the pool as memory frees up.
    void upload_loop() {
        (void)hipSetDevice(cfg.device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            while (!host_wait.empty()) {
                const JobPtr j = host_wait.front();
                if (!place_keys(j, j->host_keys.data(), hipMemcpyHostToDevice)) {
                    bool busy = false;
                    for (const auto &s : slots) busy |= s->busy;
                    if (busy || !ready.empty()) break;  // memory frees up as builds finish
                    pool.trim();                        // nothing runs: give back empty chunks
                    if (!place_keys(j, j->host_keys.data(), hipMemcpyHostToDevice)) {
                        host_wait.pop_front();
                        finish(j, RS_ERR_NOMEM, "the job's keys do not fit the HBM budget", nullptr, now());
                        cv.notify_all();
                        continue;
                    }
                }
                host_wait.pop_front();
                std::vector<uint64_t>().swap(j->host_keys);
                ready.push_back(j);
                cv.notify_all();
            }
            if (closing && host_wait.empty()) return;
            cv.wait_for(lk, std::chrono::milliseconds(20));
        }
    }
};

extern "C" {

int rs_server_create(const rs_server_config *cfg, rs_server **out) {
    if (!out) return set_error("rs_server_create: null output"), RS_ERR_ARG;
    rs_server_config c{};
    if (cfg) c = *cfg;
    if (c.slots == 0) c.slots = 4;
    if (c.slots > 64) return set_error("rs_server_create: at most 64 slots"), RS_ERR_ARG;
    if (c.hbm_budget_bytes == 0) c.hbm_budget_bytes = uint64_t(64) << 30;
    if (c.max_batch_keys == 0) c.max_batch_keys = 50000000;
    if (!(c.time_limit_s >= 0)) return set_error("rs_server_create: negative time limit"), RS_ERR_ARG;
    auto s = std::make_unique<rs_server>();
    s->cfg = c;
    s->t0 = Clock::now();
    if (hipSetDevice(c.device) != hipSuccess || hipStreamCreate(&s->upload) != hipSuccess)
        return set_error("rs_server_create: cannot use HIP device " + std::to_string(c.device)), RS_ERR_HIP;
    for (uint32_t i = 0; i < c.slots; i++) {
        auto slot = std::make_unique<rs_server::Slot>();
        if (int e = rs_ctx_create(c.device, &slot->ctx)) {
            for (auto &sl : s->slots) rs_ctx_destroy(sl->ctx);
            (void)hipStreamDestroy(s->upload);
            return e;
        }
        if (c.time_limit_s > 0) rs_ctx_set_time_limit(slot->ctx, c.time_limit_s);
        s->slots.push_back(std::move(slot));
    }
    rs_server *raw = s.release();
    for (auto &slot : raw->slots) slot->th = std::thread(&rs_server::slot_loop, raw, slot.get());
    raw->uploader = std::thread(&rs_server::upload_loop, raw);
    *out = raw;
    return RS_OK;
}

void rs_server_destroy(rs_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->closing = true;
    }
    s->cv.notify_all();
    for (auto &slot : s->slots) slot->th.join();
    s->uploader.join();
    for (auto &kv : s->jobs) rs_mphf_free(kv.second->result);
    for (auto &slot : s->slots) rs_ctx_destroy(slot->ctx);
    (void)hipSetDevice(s->cfg.device);
    (void)hipStreamDestroy(s->upload);
    delete s;
}

static int submit(rs_server *s, const uint64_t *keys, bool device_keys, size_t n, uint32_t leaf, size_t bucket,
                  int rotate, uint64_t *job_id) {
    if (!s || !keys || !job_id) return set_error("rs_submit_job: null argument"), RS_ERR_ARG;
    if (n == 0 || leaf < 2 || leaf > kMaxLeaf || bucket == 0 || n >= (uint64_t(1) << 40))
        return set_error("rs_submit_job: no keys, leaf size outside [2, 24] or zero bucket size"), RS_ERR_ARG;
    auto j = std::make_shared<Job>();
    j->leaf = leaf;
    j->bucket = bucket;
    j->rotate = rotate & (kFlagRotate | kFlagRemap64);
    j->n = n;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->closing) return set_error("rs_submit_job: server is closing"), RS_ERR_ARG;
    (void)hipSetDevice(s->cfg.device);
    j->id = s->next_id++;
    j->t_submit = s->now();
    s->st.jobs_submitted++;
    if (device_keys) {
        j->d_keys = keys;  // caller-owned, resident already
        s->ready.push_back(j);
    } else if (s->host_wait.empty() && s->place_keys(j, keys, hipMemcpyHostToDevice)) {
        s->ready.push_back(j);
    } else {
        j->host_keys.assign(keys, keys + 2 * n);  // uploaded when memory frees up
        s->host_wait.push_back(j);
    }
    s->jobs[j->id] = j;
    *job_id = j->id;
    lk.unlock();
    s->cv.notify_all();
    return RS_OK;
}

int rs_submit_job(rs_server *s, const uint64_t *keys, size_t n, uint32_t leaf_size, size_t bucket_size, int rotate,
                  uint64_t *job_id) {
    return submit(s, keys, false, n, leaf_size, bucket_size, rotate, job_id);
}

int rs_submit_job_device(rs_server *s, const uint64_t *d_keys, size_t n, uint32_t leaf_size, size_t bucket_size,
                         int rotate, uint64_t *job_id) {
    return submit(s, d_keys, true, n, leaf_size, bucket_size, rotate, job_id);
}

int rs_submit_batch(rs_server *s, const uint64_t *keys, const uint64_t *set_offsets, size_t num_sets,
                    uint32_t leaf_size, size_t bucket_size, int rotate, uint64_t *job_ids) {
    if (!s || !keys || !set_offsets || !job_ids) return set_error("rs_submit_batch: null argument"), RS_ERR_ARG;
    for (size_t i = 0; i < num_sets; i++)
        if (set_offsets[i + 1] <= set_offsets[i]) return set_error("rs_submit_batch: every set needs a key"), RS_ERR_ARG;
    for (size_t i = 0; i < num_sets; i++)
        if (int e = submit(s, keys + 2 * set_offsets[i], false, set_offsets[i + 1] - set_offsets[i], leaf_size,
                           bucket_size, rotate, &job_ids[i]))
            return e;
    return RS_OK;
}

static void fill_info(const Job &j, rs_job_info *info) {
    if (!info) return;
    info->state = j.state;
    info->status = j.status;
    info->t_submit = j.t_submit;
    info->t_start = j.t_start;
    info->t_done = j.t_done;
    info->n = j.n;
    info->leaf = j.leaf;
    info->bucket = j.bucket;
    info->batch_id = j.batch_id;
    info->batch_jobs = j.batch_jobs;
}

int rs_poll(rs_server *s, uint64_t job_id, rs_job_info *info) {
    if (!s) return set_error("rs_poll: null server"), RS_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return set_error("rs_poll: unknown job (or already handed out)"), RS_ERR_ARG;
    fill_info(*it->second, info);
    return RS_OK;
}

int rs_wait(rs_server *s, uint64_t job_id, double timeout_s, rs_mphf **out, rs_job_info *info) {
    if (!s || !out) return set_error("rs_wait: null argument"), RS_ERR_ARG;
    *out = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return set_error("rs_wait: unknown job (or already handed out)"), RS_ERR_ARG;
    const JobPtr j = it->second;
    auto done = [&] { return j->state == RS_JOB_DONE || j->state == RS_JOB_FAILED; };
    if (timeout_s < 0) s->cv.wait(lk, done);
    else if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        fill_info(*j, info);
        set_error("rs_wait: job not finished within the timeout");
        return RS_ERR_PENDING;
    }
    fill_info(*j, info);
    s->jobs.erase(j->id);
    if (j->status != RS_OK) {
        set_error("job " + std::to_string(j->id) + ": " + j->error);
        return j->status;
    }
    *out = j->result;
    j->result = nullptr;
    return RS_OK;
}

int rs_server_stats(rs_server *s, rs_serve_stats *out) {
    if (!s || !out) return set_error("rs_server_stats: null argument"), RS_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    out->committed_bytes = s->committed();
    out->key_pool_bytes = s->pool.reserved();
    out->key_pool_used_bytes = s->pool.used();
    out->jobs_waiting_for_memory = s->host_wait.size();
    out->jobs_ready = s->ready.size();
    out->uptime_s = s->now();
    return RS_OK;
}

}  // extern "C"
