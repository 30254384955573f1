// Batched build: many independent (pattern, target) jobs through ONE
// layer pipeline, the plain build's own.
//
// Every catalogue pattern this mode accepts pins nibbles 0 and 1 to 0xF, so
// the low byte of every board of such a job is 0xFF.  A board of job j is
// stored as the job key (board >> 8) | j << 56: sorting job keys groups the
// jobs (job-major) and orders each job's keys exactly as their plain values,
// and one sorted array per layer index holds that layer of every job.  The
// batched build is then the plain build's schedule over job keys -- key-set
// forward steps (sort + unique once sets outgrow their limit), unique layers
// sorted on the second stream, the final layers, the two-stream backward
// pass -- with the job-aware instantiations of the expand, solve and
// final-rate kernels, which take each key's pattern, layer count, docheck
// step and spawn weights from the job table.  Jobs with fewer layers simply
// stop contributing.  Per job and layer the results equal the plain build's
// (rates in float32).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "egtb.h"
#include "egtb_device.h"
#include "egtb_kernels.h"
#include "egtb_runtime.h"

namespace egtb {

constexpr int kMaxJobs = 256;

// Segment bounds of every job in a rotated sorted array: lower_bound of
// job << 56 for job = 0..jobs (out has jobs + 1 entries).
__global__ void k_job_bounds(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ n_ptr, int jobs,
                             uint64_t *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > jobs) return;
    const uint64_t n = *n_ptr;
    if (j == jobs && jobs == kMaxJobs) {
        out[j] = n;
        return;
    }
    const uint64_t target = (uint64_t)j << 56;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < target) lo = mid + 1;
        else hi = mid;
    }
    out[j] = lo;
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------

struct BatchResult {
    int n_jobs = 0;
    int max_steps = 0;
    std::vector<int> steps;
    Result r;                                     // global layers (job keys)
    DBuf bounds;                                  // [max_steps][n_jobs + 1]
    std::vector<std::vector<uint64_t>> job_live;  // [layer][job]
    double fwd_ms = 0, bwd_ms = 0, total_ms = 0;
};

static void batch_build(const egtb_batch_job *jobs, int n_jobs, const egtb_batch_options &bo, BatchResult &br) {
    if (n_jobs < 1 || n_jobs > kMaxJobs) throw Error(EGTB_ERR_ARG, "1..256 jobs");
    const auto t_begin = std::chrono::steady_clock::now();
    DeviceCtx &C = device_ctx(bo.device);
    hipStream_t s = C.stream;
    std::vector<JobDev> hj(n_jobs);
    int max_steps = 0;
    br.n_jobs = n_jobs;
    br.steps.resize(n_jobs);
    std::vector<uint64_t> layer0;
    for (int j = 0; j < n_jobs; ++j) {
        const egtb_batch_job &jb = jobs[j];
        if (jb.steps < 3) throw Error(EGTB_ERR_ARG, "job " + std::to_string(j) + ": steps must be at least 3");
        if (!(jb.spawn_rate4 >= 0.0 && jb.spawn_rate4 <= 1.0))
            throw Error(EGTB_ERR_ARG, "job " + std::to_string(j) + ": spawn_rate4 out of [0, 1]");
        if (!jb.seeds && jb.n_seeds) throw Error(EGTB_ERR_ARG, "job " + std::to_string(j) + ": seeds is NULL");
        hj[j].p = make_pattern(&jb.pattern, jb.target);
        // job keys need nibbles 0 and 1 pinned by every mask
        bool pinned = jb.pattern.n_masks > 0;
        for (uint32_t m = 0; m < jb.pattern.n_masks; ++m) pinned &= (jb.pattern.masks[m] & 0xFF) == 0xFF;
        if (!pinned)
            throw Error(EGTB_ERR_ARG, "job " + std::to_string(j) +
                                          ": batched builds need patterns whose masks all pin nibbles 0 and 1");
        hj[j].steps = jb.steps;
        hj[j].docheck = jb.docheck_step;
        hj[j].w4 = jb.spawn_rate4;
        hj[j].w2 = 1.0 - jb.spawn_rate4;
        max_steps = std::max(max_steps, jb.steps);
        br.steps[j] = jb.steps;
        for (size_t q = 0; q < jb.n_seeds; ++q) {
            const uint64_t k = jb.seeds[q];
            if ((k & 0xFF) != 0xFF) throw Error(EGTB_ERR_ARG, "job " + std::to_string(j) + ": seed not in pattern");
            if (j == 0 && k == 0xFF) throw Error(EGTB_ERR_ARG, "job 0: a seed's key is 0, the empty slot");
            layer0.push_back((k >> 8) | ((uint64_t)j << 56));
        }
    }
    br.max_steps = max_steps;
    DBuf d_jobs, errb;
    d_jobs.alloc(sizeof(JobDev) * n_jobs, s);
    HIPCHK(hipMemcpyAsync(d_jobs.p, hj.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice, s));
    errb.alloc(4, s);
    HIPCHK(hipMemsetAsync(errb.p, 0, 4, s));

    egtb_options o;
    std::memset(&o, 0, sizeof o);
    o.steps = max_steps;
    o.docheck_step = max_steps;  // per job, from the job table
    o.spawn_rate4 = 0.1;         // per job, from the job table
    o.success_rate_dtype = "float32";
    o.keep_raw = bo.keep_raw;
    o.device = bo.device;
    o.radix_bits = bo.radix_bits;
    o.sort_schedule = bo.sort_schedule;
    o.forward_hash_steps = bo.forward_hash_steps;
    Result &res = br.r;
    res.keep_tables = true;
    run_build_jobs(layer0.data(), layer0.size(), o, (const JobDev *)d_jobs.p, (uint32_t *)errb.p, res);

    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, errb.p, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (err & 2u) throw Error(EGTB_ERR_INTERNAL, "a canonical child left the pinned low byte");
    if (err & 4u) throw Error(EGTB_ERR_INTERNAL, "a child's job key is 0, the empty slot of key sets and hashes");

    // per-job segments of every solved layer (layer 0 keeps the seed order)
    LayerDev *M = res.meta_dev();
    br.bounds.alloc((size_t)max_steps * (n_jobs + 1) * 8, s);
    for (int i = 1; i < max_steps; ++i)
        hipLaunchKernelGGL(k_job_bounds, dim3((n_jobs + 1 + 255) / 256), dim3(256), 0, s, res.books[i].keys,
                           &M[i].n_live, n_jobs, (uint64_t *)br.bounds.p + (size_t)i * (n_jobs + 1));
    HIPCHK(hipGetLastError());
    std::vector<uint64_t> hb((size_t)max_steps * (n_jobs + 1), 0);
    HIPCHK(hipMemcpyAsync(hb.data(), br.bounds.p, hb.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    br.job_live.assign(max_steps, std::vector<uint64_t>(n_jobs, 0));
    for (int i = 1; i < max_steps; ++i) {
        hb[(size_t)i * (n_jobs + 1) + n_jobs] = res.info[i].n_live;
        for (int j = 0; j < n_jobs; ++j)
            br.job_live[i][j] = hb[(size_t)i * (n_jobs + 1) + j + 1] - hb[(size_t)i * (n_jobs + 1) + j];
    }
    HIPCHK(hipMemcpyAsync(br.bounds.p, hb.data(), hb.size() * 8, hipMemcpyHostToDevice, s));
    {
        const Book &b0 = res.books[0];
        std::vector<uint64_t> k0(res.info[0].n_live);
        if (!k0.empty()) HIPCHK(hipMemcpyAsync(k0.data(), b0.keys, k0.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (uint64_t k : k0) br.job_live[0][(int)(k >> 56)]++;
    }
    const auto t_end = std::chrono::steady_clock::now();
    br.fwd_ms = res.fwd_ms;
    br.bwd_ms = res.bwd_ms;
    br.total_ms = std::chrono::duration<double, std::milli>(t_end - t_begin).count();
}

}  // namespace egtb

// ===========================================================================
// C ABI
// ===========================================================================

using namespace egtb;

struct egtb_batch_result : BatchResult {};

template <typename F>
static int guarded_b(F &&f) {
    try {
        g_last_error.clear();
        f();
        return EGTB_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return EGTB_ERR_INTERNAL;
    }
}

extern "C" {

int egtb_batch_build(const egtb_batch_job *jobs, int n_jobs, const egtb_batch_options *options,
                     egtb_batch_result **out) {
    if (!out || !jobs || !options) return EGTB_ERR_ARG;
    *out = nullptr;
    auto br = std::make_unique<egtb_batch_result>();
    const int rc = guarded_b([&] { batch_build(jobs, n_jobs, *options, *br); });
    if (rc == EGTB_OK) *out = br.release();
    return rc;
}

int egtb_batch_timing(const egtb_batch_result *b, double *fwd_ms, double *bwd_ms, double *total_ms) {
    if (!b) return EGTB_ERR_ARG;
    if (fwd_ms) *fwd_ms = b->fwd_ms;
    if (bwd_ms) *bwd_ms = b->bwd_ms;
    if (total_ms) *total_ms = b->total_ms;
    return EGTB_OK;
}

int egtb_batch_layer_info(const egtb_batch_result *b, int layer, uint64_t *n_raw_all, uint64_t *n_live_all) {
    if (!b || layer < 0 || layer >= b->max_steps) return EGTB_ERR_ARG;
    if (n_raw_all) *n_raw_all = b->r.info[layer].n_raw;
    if (n_live_all) *n_live_all = b->r.info[layer].n_live;
    return EGTB_OK;
}

int egtb_batch_book(const egtb_batch_result *b, int job, int layer, uint64_t *keys, float *rates, size_t cap,
                    size_t *n) {
    return guarded_b([&] {
        if (!b || job < 0 || job >= b->n_jobs || layer < 0) throw Error(EGTB_ERR_ARG, "job or layer out of range");
        if (layer >= b->steps[job]) {
            if (n) *n = 0;
            return;
        }
        const Book &bk = b->r.books[layer];
        uint64_t lo, cnt = b->job_live[layer][job];
        if (layer == 0) {
            lo = 0;
            for (int j = 0; j < job; ++j) lo += b->job_live[0][j];
        } else {
            HIPCHK(hipMemcpy(&lo, (const uint64_t *)b->bounds.p + (size_t)layer * (b->n_jobs + 1) + job, 8,
                             hipMemcpyDeviceToHost));
        }
        if (n) *n = cnt;
        if (!keys && !rates) return;
        if (cap < cnt) throw Error(EGTB_ERR_ARG, "buffer too small");
        if (cnt && keys) {
            HIPCHK(hipMemcpy(keys, bk.keys + lo, cnt * 8, hipMemcpyDeviceToHost));
            for (uint64_t q = 0; q < cnt; ++q) keys[q] = (keys[q] << 8) | 0xFFull;
        }
        if (cnt && rates) HIPCHK(hipMemcpy(rates, (const float *)bk.rates + lo, cnt * 4, hipMemcpyDeviceToHost));
    });
}

void egtb_batch_free(egtb_batch_result *b) {
    if (!b) return;
    (void)hipSetDevice(b->r.dev);
    (void)hipDeviceSynchronize();
    delete b;
}

}  // extern "C"
