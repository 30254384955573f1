// Batched build: many independent (pattern, target) jobs through ONE
// layer pipeline.
//
// Every catalogue pattern this mode accepts pins nibbles 0 and 1 to 0xF, so
// the low byte of every key of such a job is 0xFF.  A key is stored rotated
// as r = (key >> 8) | (job << 56): sorting rotated keys groups the jobs
// (job-major) and orders each job's keys exactly as their plain values, and
// one sorted array per layer index holds that layer of every job.  Each
// forward step and each backward layer is then one set of kernel launches
// for all jobs: the expand / solve kernels read the job from a key's top
// byte and take that job's pattern, steps, docheck and weights from a job
// table.  Jobs with fewer layers simply stop contributing.  Per job and
// layer the results equal the plain build's (rates in float32).
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

struct JobDev {
    PatternDev p;
    int32_t steps;
    int32_t docheck;
    double w2, w4;
};

constexpr int kMaxJobs = 256;

// Largest key set of a batched forward step (128 MiB, as for plain builds).
constexpr uint64_t kBatchSetMaxSlots = (uint64_t)1 << 24;
// Largest solve hash of a batched book (as the plain builds' default).
constexpr double kBatchHashMaxBytes = 1024.0 * 1024.0 * 1024.0;

__device__ __forceinline__ uint64_t rot_key(uint64_t k, uint32_t job) { return (k >> 8) | ((uint64_t)job << 56); }
__device__ __forceinline__ uint64_t unrot_key(uint64_t r) { return (r << 8) | 0xFFull; }

// Forward step `step` over the global layer step-1.  Same rule as k_expand
// per job; children whose low byte is not 0xFF cannot be rotated and raise
// *err (the host rejects such patterns before launching).
template <bool WRITE>
__global__ __launch_bounds__(1024) void k_bexpand(const uint64_t *__restrict__ par, uint64_t n_bound,
                                                  const uint64_t *__restrict__ n_ptr,
                                                  const uint16_t *__restrict__ glut,
                                                  const JobDev *__restrict__ jobs, int step,
                                                  uint64_t *__restrict__ counts, const uint64_t *__restrict__ offs,
                                                  uint64_t *__restrict__ out2, uint64_t *__restrict__ out4,
                                                  uint64_t *__restrict__ or_and2, uint64_t *__restrict__ or_and4,
                                                  uint32_t *__restrict__ err) {
    extern __shared__ uint16_t lut[];
    __shared__ uint64_t red[4][16];
    stage_lut(lut, glut);
    const uint64_t n = n_ptr ? (*n_ptr < n_bound ? *n_ptr : n_bound) : n_bound;
    uint64_t or2 = 0, and2 = ~0ull, or4 = 0, and4 = ~0ull;
    bool bad = false, zero = false;  // zero: a rotated key equal to the solve hashes' empty slot
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t r = par[i];
        const uint32_t job = (uint32_t)(r >> 56);
        const JobDev &J = jobs[job];
        const uint64_t b = unrot_key(r);
        uint32_t c2 = 0, c4 = 0;
        uint64_t w2 = 0, w4 = 0;
        if (WRITE) {
            const uint64_t o = offs[i];
            w2 = o & 0xFFFFFFFFull;
            w4 = o >> 32;
        }
        if (step <= J.steps - 2 && !(step > J.docheck && is_success(b, J.p))) {
            uint32_t em = empty_cells(b);
            const SpawnMoves smv(b, lut);
            while (em) {
                const int k = __builtin_ctz(em);
                em &= em - 1;
#pragma unroll
                for (int s = 1; s <= 2; ++s) {
                    const uint64_t sp = b | ((uint64_t)s << (4 * k));
                    uint64_t mv[4];
                    smv.of(k, (uint64_t)s, lut, mv);
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        if (mv[d] != sp && is_pattern(mv[d], J.p)) {
                            const uint64_t c = canonical(mv[d], J.p.symm);
                            bad |= (c & 0xFF) != 0xFF;
                            const uint64_t rc = rot_key(c, job);
                            zero |= rc == 0;
                            if (s == 1) {
                                if (WRITE) out2[w2 + c2] = rc;
                                else { or2 |= rc; and2 &= rc; }
                                ++c2;
                            } else {
                                if (WRITE) out4[w4 + c4] = rc;
                                else { or4 |= rc; and4 &= rc; }
                                ++c4;
                            }
                        }
                    }
                }
            }
        }
        if (!WRITE) counts[i] = (uint64_t)c2 | ((uint64_t)c4 << 32);
    }
    if (bad) atomicOr(err, 2u);
    if (zero) atomicOr(err, 4u);
    if (!WRITE) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            or2 |= __shfl_xor(or2, off, 64);
            and2 &= __shfl_xor(and2, off, 64);
            or4 |= __shfl_xor(or4, off, 64);
            and4 &= __shfl_xor(and4, off, 64);
        }
        const uint32_t wave = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            red[0][wave] = or2;
            red[1][wave] = and2;
            red[2][wave] = or4;
            red[3][wave] = and4;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t w = 1; w < (blockDim.x >> 6); ++w) {
                or2 |= red[0][w];
                and2 &= red[1][w];
                or4 |= red[2][w];
                and4 &= red[3][w];
            }
            atomicOr((unsigned long long *)&or_and2[0], (unsigned long long)or2);
            atomicAnd((unsigned long long *)&or_and2[1], (unsigned long long)and2);
            atomicOr((unsigned long long *)&or_and4[0], (unsigned long long)or4);
            atomicAnd((unsigned long long *)&or_and4[1], (unsigned long long)and4);
        }
    }
}

// Backward layer `layer` for every job: final layers (success -> 1, else 0)
// for jobs whose last two layers these are, the recurrence of k_solve
// otherwise, looking up rotated keys of the same job in the global books.
__global__ __launch_bounds__(1024) void k_bsolve(const uint64_t *__restrict__ keys, uint64_t n,
                                                 const uint16_t *__restrict__ glut, const JobDev *__restrict__ jobs,
                                                 int layer, TableDev<float> t1, TableDev<float> t2, HashDev h1,
                                                 HashDev h2, float *__restrict__ rates) {
    extern __shared__ uint16_t lut[];
    stage_lut(lut, glut);
    // hash probes when both successor layers have a hash (rotated keys; the
    // forward pass rejects a rotated key of 0), else bucket-index searches
    const bool use_hash = h1.slots != nullptr && h2.slots != nullptr;
    const uint64_t n1 = t1.n_ptr ? *t1.n_ptr : t1.n, n2 = t2.n_ptr ? *t2.n_ptr : t2.n;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t r = keys[i];
        const uint32_t job = (uint32_t)(r >> 56);
        const JobDev &J = jobs[job];
        const uint64_t b = unrot_key(r);
        if (layer >= J.steps - 2) {
            rates[i] = is_success(b, J.p) ? 1.0f : 0.0f;
            continue;
        }
        if (layer > J.docheck && is_success(b, J.p)) {
            rates[i] = 1.0f;
            continue;
        }
        double sp = 0.0;
        uint32_t em = empty_cells(b);
        const int e = __popc(em);
        const SpawnMoves smv(b, lut);
        while (em) {
            const int k = __builtin_ctz(em);
            em &= em - 1;
            uint64_t key[8];
            uint32_t lo[8], hi[8];
            uint32_t live = 0;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const uint64_t spb = b | ((uint64_t)(s + 1) << (4 * k));
                uint64_t mv[4];
                smv.of(k, (uint64_t)(s + 1), lut, mv);
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const int j = s * 4 + d;
                    key[j] = rot_key(canonical(mv[d], J.p.symm), job);
                    lo[j] = hi[j] = 0;
                    const uint64_t nn = s == 0 ? n1 : n2;
                    if (mv[d] != spb && is_pattern(mv[d], J.p) && nn) {
                        if (!use_hash) {
                            const TableDev<float> &t = s == 0 ? t1 : t2;
                            const uint64_t bk = key[j] >> (64 - t.bits);
                            lo[j] = t.idx[bk];
                            hi[j] = t.idx[bk + 1];
                        }
                        live |= 1u << j;
                    }
                }
            }
            float best[2] = {0.0f, 0.0f};
            if (use_hash) {
                uint64_t slot[8];
                float r[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    slot[j] = hash_mix(key[j]) & (j < 4 ? h1.mask : h2.mask);
                    r[j] = 0.0f;
                }
                uint32_t active = live;
                while (active) {
                    ulonglong2 ent[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j)
                        if ((active >> j) & 1u) ent[j] = (j < 4 ? h1.slots : h2.slots)[slot[j]];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if ((active >> j) & 1u) {
                            if (ent[j].x == key[j]) {
                                r[j] = rate_from_bits<float>(ent[j].y);
                                active &= ~(1u << j);
                            } else if (ent[j].x == 0) {
                                active &= ~(1u << j);  // absent: rate 0
                            } else {
                                slot[j] = (slot[j] + 1) & (j < 4 ? h1.mask : h2.mask);
                            }
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if ((live >> j) & 1u) best[j >> 2] = best[j >> 2] < r[j] ? r[j] : best[j >> 2];
            } else {
                uint32_t active = live;
                while (active) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (lo[j] < hi[j]) {
                            const uint32_t mid = (lo[j] + hi[j]) >> 1;
                            const uint64_t km = (j < 4 ? t1.keys : t2.keys)[mid];
                            if (km < key[j]) lo[j] = mid + 1;
                            else hi[j] = mid;
                        }
                    }
                    active = 0;
#pragma unroll
                    for (int j = 0; j < 8; ++j) active |= (lo[j] < hi[j] ? 1u : 0u) << j;
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if ((live >> j) & 1u) {
                        const TableDev<float> &t = j < 4 ? t1 : t2;
                        const uint64_t nn = j < 4 ? n1 : n2;
                        const float v = (lo[j] < nn && t.keys[lo[j]] == key[j]) ? t.rates[lo[j]] : 0.0f;
                        best[j >> 2] = best[j >> 2] < v ? v : best[j >> 2];
                    }
                }
            }
            sp = __fma_rn((double)best[0], J.w2, sp);
            sp = __fma_rn((double)best[1], J.w4, sp);
        }
        rates[i] = e > 0 ? (float)(sp / (double)e) : 0.0f;
    }
}

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

static void configure_batch_kernels() {
    const int lds = kLutEntries * sizeof(uint16_t);
    (void)hipFuncSetAttribute((const void *)k_bexpand<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_bexpand<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void *)k_bsolve, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------

struct BatchResult {
    int n_jobs = 0;
    int max_steps = 0;
    std::vector<int> steps;
    std::vector<std::vector<uint64_t>> seeds;  // per job, catalogue order (layer 0)
    Result r;                                  // global layers (rotated keys)
    DBuf bounds;                               // [max_steps][n_jobs + 1]
    std::vector<std::vector<uint64_t>> job_live;  // [layer][job]
    std::vector<std::vector<uint64_t>> job_raw;   // [layer][job]
    double fwd_ms = 0, bwd_ms = 0, total_ms = 0;
};

static uint32_t bindex_bits(uint64_t n) {
    uint32_t b = 1;
    while (b < 24 && (1ull << (b + 3)) < n) ++b;
    return b;
}

static void batch_build(const egtb_batch_job *jobs, int n_jobs, const egtb_batch_options &bo, BatchResult &br) {
    if (n_jobs < 1 || n_jobs > kMaxJobs) throw Error(EGTB_ERR_ARG, "1..256 jobs");
    DeviceCtx &C = device_ctx(bo.device);
    configure_batch_kernels();
    hipStream_t s = C.stream;
    std::vector<JobDev> hj(n_jobs);
    int max_steps = 0;
    br.n_jobs = n_jobs;
    br.steps.resize(n_jobs);
    br.seeds.resize(n_jobs);
    std::vector<uint64_t> layer0;
    for (int j = 0; j < n_jobs; ++j) {
        const egtb_batch_job &jb = jobs[j];
        if (jb.steps < 3) throw Error(EGTB_ERR_ARG, "job " + std::to_string(j) + ": steps must be at least 3");
        hj[j].p = make_pattern(&jb.pattern, jb.target);
        // the rotation needs nibbles 0 and 1 pinned by every mask
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
            if (j == 0 && k == 0xFF) throw Error(EGTB_ERR_ARG, "job 0: a seed rotates to key 0, the empty slot");
            br.seeds[j].push_back(k);
            layer0.push_back((k >> 8) | ((uint64_t)j << 56));
        }
    }
    br.max_steps = max_steps;
    Result &res = br.r;
    res.dev = bo.device;
    res.steps = max_steps;
    res.rate_size = 4;
    res.rate_is_float = true;
    res.raw.assign(max_steps, nullptr);
    res.books.assign(max_steps, Book{});
    res.info.assign(max_steps, LayerInfo{});
    DBuf d_jobs, cand[2], sort_tmp, counts, offs, rates, errb;
    d_jobs.alloc(sizeof(JobDev) * n_jobs, s);
    HIPCHK(hipMemcpyAsync(d_jobs.p, hj.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice, s));
    errb.alloc(4, s);
    HIPCHK(hipMemsetAsync(errb.p, 0, 4, s));
    SortWork w;
    HostStage meta_host;
    const auto t_begin = std::chrono::steady_clock::now();

    meta_host.ensure((size_t)(max_steps + 1) * sizeof(LayerDev));
    LayerDev *H = (LayerDev *)meta_host.p;
    for (int i = 0; i <= max_steps; ++i) {
        std::memset(&H[i], 0, sizeof(LayerDev));
        H[i].and_all = ~0ull;
    }
    H[0].n_raw = layer0.size();
    res.meta.alloc((size_t)(max_steps + 1) * sizeof(LayerDev), s);
    HIPCHK(hipMemcpyAsync(res.meta.p, H, (size_t)(max_steps + 1) * sizeof(LayerDev), hipMemcpyHostToDevice, s));
    LayerDev *M = res.meta_dev();
    res.raw[0] = (uint64_t *)res.raw_arena.alloc(layer0.size() * 8);
    HIPCHK(hipMemcpyAsync(res.raw[0], layer0.data(), layer0.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));  // layer0 is a host vector
    res.info[0].n_raw = layer0.size();

    const size_t lds = kLutEntries * sizeof(uint16_t);
    auto grid_for = [&](uint64_t n) {
        const uint64_t want = (n + C.shape.threads - 1) / C.shape.threads;
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, C.shape.blocks));
    };
    auto launch_count = [&](int step, uint64_t bound) {
        counts.ensure(bound * 8 + 8, s);
        offs.ensure(bound * 8 + 8, s);
        w.partial.ensure(scan_partials_needed(bound) * 8 + 64, s);
        if (bound) {
            Timer t(s, "batch_expand_count", 8.0 * (double)bound);
            hipLaunchKernelGGL(k_bexpand<false>, dim3(grid_for(bound)), dim3(C.shape.threads), lds, s,
                               res.raw[step - 1], bound, &M[step - 1].n_raw, C.lut, (const JobDev *)d_jobs.p, step,
                               (uint64_t *)counts.p, nullptr, nullptr, nullptr, &M[step].or_all,
                               &M[step + 1].or_all, (uint32_t *)errb.p);
        }
        launch_scan_u64(s, (const uint64_t *)counts.p, bound, &M[step - 1].n_raw, (uint64_t *)offs.p,
                        (uint64_t *)w.partial.p, &M[step].tot);
    };

    // Layer L from its candidates keys[0..n_c): into a key set of at least
    // 2 n_c slots (it cannot fill), extracted unsorted, then only the
    // unique keys sorted into the raw arena; sets above kBatchSetMaxSlots
    // sort and unique the candidates instead. The raw layer and M[L].n_raw
    // are the same either way.
    DBuf kset, uniq;
    auto dedup_layer = [&](int L, uint64_t *keys, uint64_t n_c) {
        uint64_t cap = 1024;
        while (cap < 2 * n_c) cap <<= 1;
        if (n_c >= 2 && cap <= kBatchSetMaxSlots && bo.sort_schedule == 0) {
            if (kset.bytes < cap * 8) {
                kset.alloc(cap * 8, s);
                HIPCHK(hipMemsetAsync(kset.p, 0, kset.bytes, s));
            }
            uniq.ensure(n_c * 8, s);
            {
                Timer t(s, "batch_set", 8.0 * (double)n_c + 16.0 * (double)cap);
                launch_set_insert(s, keys, n_c, (uint64_t *)kset.p, cap);
                HIPCHK(hipMemsetAsync(&M[L].n_raw, 0, 8, s));
                launch_set_extract(s, (uint64_t *)kset.p, cap, (uint64_t *)uniq.p, &M[L].n_raw);
            }
            HIPCHK(hipMemcpyAsync(&H[L].n_raw, &M[L].n_raw, 8, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            const uint64_t n = H[L].n_raw;
            res.raw[L] = (uint64_t *)res.raw_arena.alloc(n * 8);
            Timer t(s, "batch_sort", 16.0 * (double)n);
            uint64_t *out = radix_sort(s, (uint64_t *)uniq.p, res.raw[L], n, H[L].or_all, H[L].and_all,
                                       bo.radix_bits, w, 0);
            if (out != res.raw[L] && n) HIPCHK(hipMemcpyAsync(res.raw[L], out, n * 8, hipMemcpyDeviceToDevice, s));
            return;
        }
        uint64_t *sorted;
        {
            Timer t(s, "batch_sort", 16.0 * (double)n_c);
            sorted = radix_sort(s, keys, (uint64_t *)sort_tmp.p, n_c, H[L].or_all, H[L].and_all, bo.radix_bits, w,
                                bo.sort_schedule);
        }
        res.raw[L] = (uint64_t *)res.raw_arena.alloc(n_c * 8);
        {
            Timer t(s, "batch_unique", 16.0 * (double)n_c);
            unique_into(s, sorted, n_c, res.raw[L], &M[L].n_raw, w, bo.sort_schedule);
        }
    };

    uint64_t c4_prev = 0;
    launch_count(1, layer0.size());
    for (int i = 1; i <= max_steps - 2; ++i) {
        HIPCHK(hipMemcpyAsync(H + i - 1, M + i - 1, 2 * sizeof(LayerDev), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const uint64_t n_par = H[i - 1].n_raw;
        if (i > 1) res.raw_arena.trim_last(res.raw[i - 1], n_par * 8);
        res.info[i - 1].n_raw = n_par;
        const uint64_t c2 = H[i].tot & 0xFFFFFFFFull, c4 = H[i].tot >> 32;
        res.info[i].c2 = c2;
        res.info[i].c4 = c4;
        const uint64_t n_c = c4_prev + c2;
        DBuf &cur = cand[i & 1], &nxt = cand[(i + 1) & 1];
        if (cur.bytes < n_c * 8) {
            DBuf g;
            g.alloc(n_c * 8 + n_c * 4 + 256, s);
            if (c4_prev) HIPCHK(hipMemcpyAsync(g.p, cur.p, c4_prev * 8, hipMemcpyDeviceToDevice, s));
            cur = std::move(g);
        }
        nxt.ensure(c4 * 8, s);
        sort_tmp.ensure(n_c * 8, s);
        if (n_par && (c2 || c4)) {
            Timer t(s, "batch_expand", 8.0 * (double)n_par + 8.0 * (double)(c2 + c4));
            hipLaunchKernelGGL(k_bexpand<true>, dim3(grid_for(n_par)), dim3(C.shape.threads), lds, s, res.raw[i - 1],
                               n_par, nullptr, C.lut, (const JobDev *)d_jobs.p, i, nullptr, (const uint64_t *)offs.p,
                               (uint64_t *)cur.p + c4_prev, (uint64_t *)nxt.p, nullptr, nullptr, (uint32_t *)errb.p);
        }
        dedup_layer(i, (uint64_t *)cur.p, n_c);
        if (i + 1 <= max_steps - 2) launch_count(i + 1, n_c);
        HIPCHK(hipGetLastError());
        c4_prev = c4;
    }
    {
        const int L = max_steps - 1;
        HIPCHK(hipMemcpyAsync(H + L - 1, M + L - 1, 2 * sizeof(LayerDev), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        res.raw_arena.trim_last(res.raw[L - 1], H[L - 1].n_raw * 8);
        res.info[L - 1].n_raw = H[L - 1].n_raw;
        const uint64_t n_c = c4_prev;
        DBuf &cur = cand[L & 1];
        sort_tmp.ensure(n_c * 8, s);
        dedup_layer(L, (uint64_t *)cur.p, n_c);
        HIPCHK(hipMemcpyAsync(H + L, M + L, sizeof(LayerDev), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        res.raw_arena.trim_last(res.raw[L], H[L].n_raw * 8);
        res.info[L].n_raw = H[L].n_raw;
    }
    cand[0].release();
    cand[1].release();
    sort_tmp.release();
    kset.release();
    uniq.release();
    counts.release();
    offs.release();
    const auto t_fwd = std::chrono::steady_clock::now();

    // ---- backward, all jobs per layer; sizes stay on the device
    DBuf hring[3];  // solve hashes of the books, layer i in slot i % 3
    uint64_t max_n = 0;
    for (int i = 0; i < max_steps; ++i) max_n = std::max(max_n, res.info[i].n_raw);
    rates.ensure(max_n * 4 + 8, s);
    br.bounds.alloc((size_t)max_steps * (n_jobs + 1) * 8, s);
    for (int i = max_steps - 1; i >= 0; --i) {
        const uint64_t n = res.info[i].n_raw;
        TableDev<float> t1{}, t2{};
        if (i + 1 < max_steps) {
            const Book &b1 = res.books[i + 1];
            t1 = TableDev<float>{b1.keys, (const float *)b1.rates, b1.idx, b1.n_dev, 0, b1.bits, 1};
        }
        if (i + 2 < max_steps) {
            const Book &b2 = res.books[i + 2];
            t2 = TableDev<float>{b2.keys, (const float *)b2.rates, b2.idx, b2.n_dev, 0, b2.bits, 1};
        }
        HashDev h1{}, h2{};
        if (i + 2 < max_steps) {
            h1 = res.books[i + 1].h;
            h2 = res.books[i + 2].h;
        }
        if (n) {
            Timer t(s, "batch_solve", (double)n * 12.0);
            hipLaunchKernelGGL(k_bsolve, dim3(grid_for(n)), dim3(C.shape.threads), lds, s, res.raw[i], n, C.lut,
                               (const JobDev *)d_jobs.p, i, t1, t2, h1, h2, (float *)rates.p);
        }
        // compact + index (the same kernels as the plain build)
        Book &bk = res.books[i];
        bk.n_dev = &M[i].n_live;
        bk.keys = (uint64_t *)res.book_arena.alloc(n * 8);
        bk.rates = res.book_arena.alloc(n * 4);
        if (n) w.ensure_lookback(s, 0, 0, compact_tiles(n));
        // the compaction fills layer i's solve hash (read by the solves of
        // layers i-1 and i-2): ring slot i % 3, whose previous table (layer
        // i+3) was last read by the solve of layer i+1
        ulonglong2 *hs = nullptr;
        uint64_t hcap = 0;
        if (i > 0) {
            hcap = hash_capacity(n);
            if ((double)hcap * sizeof(ulonglong2) <= kBatchHashMaxBytes) {
                hring[i % 3].ensure(hcap * sizeof(ulonglong2), s);
                hs = (ulonglong2 *)hring[i % 3].p;
            }
        }
        bk.h = HashDev{};
        if (hs) {  // rotated keys span every job: the wide layout
            bk.h.slots = hs;
            bk.h.mask = hcap - 1;
        }
        launch_compact<float>(s, res.raw[i], (const float *)rates.p, n, bk.keys, (float *)bk.rates,
                              (unsigned long long *)&M[i].maxbits, &M[i].n_live, w.lw, 0.0f, bk.h);
        bk.bits = bindex_bits(n);
        bk.idx = (uint32_t *)res.book_arena.alloc(((1ull << bk.bits) + 1) * 4);
        launch_bucket_index(s, bk.keys, n, bk.n_dev, bk.bits, bk.idx);
        if (i > 0)
            hipLaunchKernelGGL(k_job_bounds, dim3((n_jobs + 1 + 255) / 256), dim3(256), 0, s, bk.keys, bk.n_dev,
                               n_jobs, (uint64_t *)br.bounds.p + (size_t)i * (n_jobs + 1));
        HIPCHK(hipGetLastError());
    }
    for (int i = 0; i < max_steps; ++i) res.books[i].h = HashDev{};  // the ring is freed with this scope
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, errb.p, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(H, M, (size_t)max_steps * sizeof(LayerDev), hipMemcpyDeviceToHost, s));
    std::vector<uint64_t> hb((size_t)max_steps * (n_jobs + 1));
    HIPCHK(hipMemcpyAsync(hb.data(), br.bounds.p, hb.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (err & 2u) throw Error(EGTB_ERR_INTERNAL, "a canonical child left the pinned low byte");
    if (err & 4u) throw Error(EGTB_ERR_INTERNAL, "a rotated key is 0, which the solve hashes cannot hold");
    if (w.read_error(s)) throw Error(EGTB_ERR_INTERNAL, "a look-back kernel exceeded its spin limit");
    const auto t_end = std::chrono::steady_clock::now();
    for (int i = 0; i < max_steps; ++i) {
        res.info[i].n_live = H[i].n_live;
        res.books[i].n = H[i].n_live;
    }
    // per-job live counts (layer 0 from its seed order, others from bounds)
    br.job_live.assign(max_steps, std::vector<uint64_t>(n_jobs, 0));
    for (int i = 1; i < max_steps; ++i)
        for (int j = 0; j < n_jobs; ++j) {
            const uint64_t a = hb[(size_t)i * (n_jobs + 1) + j];
            const uint64_t b = j + 1 < n_jobs ? hb[(size_t)i * (n_jobs + 1) + j + 1] : H[i].n_live;
            br.job_live[i][j] = b - a;
        }
    std::vector<uint64_t> hb_fixed = hb;
    for (int i = 1; i < max_steps; ++i) hb_fixed[(size_t)i * (n_jobs + 1) + n_jobs] = H[i].n_live;
    HIPCHK(hipMemcpy(br.bounds.p, hb_fixed.data(), hb_fixed.size() * 8, hipMemcpyHostToDevice));
    if (!bo.keep_raw) {
        res.raw_arena.release();
        res.raw.assign(max_steps, nullptr);
    }
    br.fwd_ms = std::chrono::duration<double, std::milli>(t_fwd - t_begin).count();
    br.bwd_ms = std::chrono::duration<double, std::milli>(t_end - t_fwd).count();
    br.total_ms = std::chrono::duration<double, std::milli>(t_end - t_begin).count();
    // layer 0: split by job on the host (tiny, seed order)
    {
        const Book &b0 = res.books[0];
        std::vector<uint64_t> k0(b0.n);
        if (b0.n) HIPCHK(hipMemcpy(k0.data(), b0.keys, b0.n * 8, hipMemcpyDeviceToHost));
        for (uint64_t q = 0; q < b0.n; ++q) br.job_live[0][(int)(k0[q] >> 56)]++;
    }
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
