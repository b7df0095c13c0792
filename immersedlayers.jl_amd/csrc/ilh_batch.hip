// ilh_batch.hip -- many independent Dirichlet repeat solves (the batched form
// of test/literate/problems.jl:259-298) in ONE launch per pass over all
// problems:
//   1. surface step   s = M+ f+ (+ M- f-), amp = [s; f+ - f-]   (4 rows a tile)
//   2. row pass       the merged surface sources' row pairs -> T1 (per problem)
//   3. column pass    forward column FFT, LGF spectrum, inverse -> T2
//   4. inverse rows   T2 -> the grid field f
// A pass's tiles (problem, logical work-group) from every problem form one
// tile list.  The list is cut into 8 segments, one per XCD, with whole
// problems balanced over the segments by estimated cost (largest first), so
// that a problem's tiles share one L2.  Two schedules:
//   dispatch  one work-group per tile; the hardware hands work-groups to the
//             XCDs round-robin, renumbered so XCD x runs segment x;
//   queue     a persistent grid of `mult` work-groups per CU pulls tiles from
//             its XCD's queue (an atomic head per segment) and, once that is
//             empty, from the other XCDs' queues.  No work-group ever waits
//             for another; the last one out resets the heads for the next
//             launch (graph replays included).
// Transforms per work-group (256 or 512 threads: NB = threads / (P/8)
// transforms of each problem's length P), the schedule and the grid size of
// every pass, and whether the row pass stages amplitudes in LDS, are chosen
// by the autotuner (fused_tune): it times each candidate of each pass on the
// batch itself and keeps the fastest.  Every transform runs the same code in
// the same order as the single-problem kernels, so a batched result equals
// its single solve bit for bit.
#include "ilh_fft_dev.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>

namespace ilh {

namespace {

constexpr unsigned NO_TILE = 0xFFFFFFFFu;
enum { SCHED_DISPATCH = 0, SCHED_QUEUE = 1 };
enum { PASS_AMP = 0, PASS_ROWS = 1, PASS_COLS = 2, PASS_INV = 3 };

// one pass's tile list: 8 segments of `seg` codes (problem << 16 | block),
// segment x holding len[x] tiles then NO_TILE padding; the queue heads
struct alignas(128) TileQueue {
    int head[8];
    int done;
    int seg;
    int len[8];
};

// The tiles of one launch.  `body(code, tid)` runs one tile with the
// work-group's thread index tid.  In the queue loop tid is re-read through an
// opaque move each iteration: otherwise the compiler hoists every
// thread-index-derived address of the (fully unrolled) transforms out of the
// loop and holds them all in registers across it (hundreds of VGPRs, spills).
template <int SCHED, class Body>
__device__ __forceinline__ void run_tiles(const unsigned* __restrict__ tiles, TileQueue* q, Body&& body)
{
    if constexpr (SCHED == SCHED_DISPATCH) {
        const unsigned code = tiles[xcd_block_of(blockIdx.x, gridDim.x)];
        if (code != NO_TILE) body(code, (int)threadIdx.x);
    } else {
        __shared__ int s_tile;
        const int home = blockIdx.x & 7;
        const int seg = q->seg;
        int qi = 0;   // queues tried so far (thread 0)
        for (;;) {
            __syncthreads();   // the previous tile's LDS use and the read of s_tile are over
            if (threadIdx.x == 0) {
                int t = -1;
                for (; qi < 8; ++qi) {
                    const int x = (home + qi) & 7;
                    const int len = q->len[x];
                    if (len == 0) continue;
                    const int k = atomicAdd(&q->head[x], 1);
                    if (k < len) {
                        t = x * seg + k;
                        break;
                    }
                }
                s_tile = t;
            }
            __syncthreads();
            const int t = __builtin_amdgcn_readfirstlane(s_tile);
            if (t < 0) break;
            int tid;
            asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
            body(tiles[t], tid);
        }
        // the last work-group out resets the queue for the next launch
        if (threadIdx.x == 0 && atomicAdd(&q->done, 1) == (int)gridDim.x - 1) {
            for (int x = 0; x < 8; ++x) atomicExch(&q->head[x], 0);
            atomicExch(&q->done, 0);
        }
    }
}

// ----------------------------------------------------------------- kernels
// Pass 1: the Dirichlet surface step, 4 rows a tile, one wave a row (the
// arithmetic of the single solve's k_dir_amp).
template <int SCHED>
__global__ void __launch_bounds__(256) k_batch_amp(const BatchProb* __restrict__ pr, const unsigned* __restrict__ tiles,
                                                   TileQueue* q)
{
    run_tiles<SCHED>(tiles, q, [&](unsigned code, int tid) {
        const BatchProb& B = pr[code >> 16];
        const int n = B.n;
        const int i = (int)(code & 0xFFFF) * 4 + (tid >> 6), lane = tid & 63;
        if (i >= n) return;
        double s = wave_dot<1>(B.Mp + (size_t)B.ld * i, B.fp, n, lane);
        if (B.fm) s += wave_dot<1>(B.Mm + (size_t)B.ld * i, B.fm, n, lane);
        if (lane == 0) {
            B.amp[i] = s;
            B.amp[n + i] = B.fp[i] - (B.fm ? B.fm[i] : 0.0);
            B.s[i] = s;
        }
    });
}

#define ILH_BATCH_FOR_P(PV, BODY)                      \
    switch (PV) {                                      \
    case 128: { constexpr int PP = 128; BODY; } break;   \
    case 256: { constexpr int PP = 256; BODY; } break;   \
    case 512: { constexpr int PP = 512; BODY; } break;   \
    case 1024: { constexpr int PP = 1024; BODY; } break; \
    case 2048: { constexpr int PP = 2048; BODY; } break; \
    default: break;                                    \
    }

// Pass 2: row pairs of the merged surface sources (amplitudes staged in LDS
// when 2n <= amp_cap).
template <int TPB, int SCHED>
__global__ void __launch_bounds__(TPB) k_batch_rows(const BatchProb* __restrict__ pr,
                                                    const unsigned* __restrict__ tiles, TileQueue* q, int amp_cap)
{
    extern __shared__ double2 lds_raw[];
    run_tiles<SCHED>(tiles, q, [&](unsigned code, int tid) {
        const BatchProb& B = pr[code >> 16];
        const int blk = code & 0xFFFF;
        RowSrc S{};
        S.mode = ROW_TERMS;
        S.nx = B.nx;
        S.ny = B.ny;
        S.pinfo = B.pinfo;
        S.cells = B.cells;
        S.cai = B.cai;
        S.cw = B.cw;
        S.amp = B.amp;
        S.amp_len = 2 * B.n;
        S.amp_lds = 2 * B.n <= amp_cap;
        // a live count on the device: tiles past it have no pairs
        const int np = B.npairs_dev ? *B.npairs_dev : B.npairs;
        ILH_BATCH_FOR_P(B.Px, if (blk * (TPB / (PP / V)) >= np) return;
                        (rows_fwd_body<PP, TPB / (PP / V), double, true>(S, B.twx, np, B.ny, B.T1, blk, lds_raw,
                                                                          tid)));
    });
}

// Pass 3: spectrum columns (forward, LGF spectrum, inverse)
template <int TPB, int SCHED>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4)))
k_batch_cols(const BatchProb* __restrict__ pr, const unsigned* __restrict__ tiles, TileQueue* q)
{
    extern __shared__ double2 lds_raw[];
    run_tiles<SCHED>(tiles, q, [&](unsigned code, int tid) {
        const BatchProb& B = pr[code >> 16];
        const int blk = code & 0xFFFF;
        ILH_BATCH_FOR_P(B.Py, (cols_body<PP, TPB / (PP / V), 0, double, double>(
                                  B.T1, B.slot, B.twy, B.twy, B.Px / 2, B.ny, 0, B.spec, 1.0, B.T2, nullptr, blk,
                                  lds_raw, tid)));
    });
}

// Pass 4: inverse row pairs into the grid field
template <int TPB, int SCHED>
__global__ void __launch_bounds__(TPB) k_batch_inv(const BatchProb* __restrict__ pr, const unsigned* __restrict__ tiles,
                                                   TileQueue* q)
{
    extern __shared__ double2 lds_raw[];
    run_tiles<SCHED>(tiles, q, [&](unsigned code, int tid) {
        const BatchProb& B = pr[code >> 16];
        const int blk = code & 0xFFFF;
        ILH_BATCH_FOR_P(B.Px, (rows_inv_body<PP, TPB / (PP / V), double, double>(B.T2, B.twx, B.nx, B.ny, B.out, blk,
                                                                                 lds_raw, tid)));
    });
}

// ------------------------------------------------------------------- host
constexpr int NPASS = 4;
constexpr int TPBS[2] = {256, 512};

// the pass's tile list for one work-group size (amp: 256 only)
struct TileList {
    size_t off = 0;     // first code in FusedBatch::tiles
    int seg = 0;
    int len[8] = {};
    int ntiles = 0;
};

int log2i(int P)
{
    int l = 0;
    while ((1 << l) < P) ++l;
    return l;
}

void set_lds_batch(const void* fn, size_t bytes)
{
    static std::map<const void*, size_t> lim;
    auto it = lim.find(fn);
    if (it != lim.end() && it->second >= bytes) return;
    ILH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    lim[fn] = bytes;
}

}  // namespace

struct FusedBatch {
    ilh_ctx* ctx = nullptr;
    std::vector<BatchProb> probs;
    DevBuf d_probs, d_tiles, d_queues;
    std::vector<unsigned> h_tiles;
    TileList lists[NPASS][2];   // [pass][work-group size index]
    FusedCfg cfg;
    float tuned_ms[NPASS] = {};
    int max_amp = 0;            // largest amplitude count 2n
};

bool fused_supported(int Px, int Py)
{
    auto ok = [](int P) { return P >= 128 && P <= 2048 && (P & (P - 1)) == 0; };
    return ok(Px) && ok(Py);
}

FusedCfg fused_default()
{
    FusedCfg c;
    for (int p = 0; p < NPASS; ++p) {
        c.sched[p] = SCHED_DISPATCH;
        c.mult[p] = 2;
        c.tpb[p] = 256;
    }
    c.amp_cap = 4096;
    return c;
}

FusedBatch* fused_create(ilh_ctx* ctx, const std::vector<BatchProb>& probs)
{
    if (probs.empty()) fail(ILH_ERR_ARG, "fused batch of no problems");
    if (probs.size() > 65535) fail(ILH_ERR_ARG, "fused batch of more than 65535 problems");
    std::unique_ptr<FusedBatch> fb(new FusedBatch());
    fb->ctx = ctx;
    fb->probs = probs;
    fb->cfg = fused_default();
    const int np = (int)probs.size();
    for (const BatchProb& b : probs) {
        if (b.n <= 0) fail(ILH_ERR_ARG, "fused batch problem without surface points");
        if (!fused_supported(b.Px, b.Py)) fail(ILH_ERR_ARG, "fused batch problem with an unsupported FFT period");
        if ((b.n + 3) / 4 > 65535) fail(ILH_ERR_ARG, "fused batch problem with too many surface points");
        fb->max_amp = std::max(fb->max_amp, 2 * b.n);
    }
    // tiles of every (pass, work-group size): per problem a block count and
    // a per-block cost; problems balanced over 8 segments, largest first
    for (int pass = 0; pass < NPASS; ++pass)
        for (int w = 0; w < 2; ++w) {
            if (pass == PASS_AMP && w == 1) continue;
            const int tpb = TPBS[w];
            std::vector<int> nblk(np);
            std::vector<double> bcost(np);
            for (int i = 0; i < np; ++i) {
                const BatchProb& b = probs[i];
                if (pass == PASS_AMP) {
                    nblk[i] = (b.n + 3) / 4;
                    bcost[i] = 4.0 * b.n;
                } else {
                    const int P = pass == PASS_COLS ? b.Py : b.Px;
                    const int nb = tpb / (P / V);
                    const int units = pass == PASS_ROWS ? b.npairs : pass == PASS_COLS ? b.Px / 2 : (b.ny + 1) / 2;
                    nblk[i] = (units + nb - 1) / nb;
                    bcost[i] = (double)nb * P * log2i(P);
                }
                if (nblk[i] > 65535) fail(ILH_ERR_ARG, "fused batch problem too large");
            }
            std::vector<int> order(np);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](int a, int b) { return nblk[a] * bcost[a] > nblk[b] * bcost[b]; });
            double load[8] = {};
            std::vector<int> seg_of(np);
            for (int i : order) {
                const int x = (int)(std::min_element(load, load + 8) - load);
                seg_of[i] = x;
                load[x] += nblk[i] * bcost[i];
            }
            TileList& L = fb->lists[pass][w];
            std::vector<std::vector<unsigned>> segs(8);
            // within a segment: costlier tiles first
            std::vector<int> by_tile(np);
            std::iota(by_tile.begin(), by_tile.end(), 0);
            std::stable_sort(by_tile.begin(), by_tile.end(), [&](int a, int b) { return bcost[a] > bcost[b]; });
            for (int i : by_tile)
                for (int k = 0; k < nblk[i]; ++k) segs[seg_of[i]].push_back(((unsigned)i << 16) | (unsigned)k);
            int seg = 1;
            for (int x = 0; x < 8; ++x) seg = std::max(seg, (int)segs[x].size());
            L.off = fb->h_tiles.size();
            L.seg = seg;
            L.ntiles = 0;
            for (int x = 0; x < 8; ++x) {
                L.len[x] = (int)segs[x].size();
                L.ntiles += L.len[x];
                segs[x].resize(seg, NO_TILE);
                fb->h_tiles.insert(fb->h_tiles.end(), segs[x].begin(), segs[x].end());
            }
        }
    hipStream_t st = ctx->stream;
    fb->d_probs.ensure(sizeof(BatchProb) * probs.size());
    fb->d_tiles.ensure(sizeof(unsigned) * fb->h_tiles.size());
    ILH_HIP(hipMemcpyAsync(fb->d_probs.p, fb->probs.data(), sizeof(BatchProb) * probs.size(), hipMemcpyHostToDevice,
                           st));
    ILH_HIP(hipMemcpyAsync(fb->d_tiles.p, fb->h_tiles.data(), sizeof(unsigned) * fb->h_tiles.size(),
                           hipMemcpyHostToDevice, st));
    // one queue per (pass, work-group size): heads and done at zero
    std::vector<TileQueue> hq(NPASS * 2);
    for (int pass = 0; pass < NPASS; ++pass)
        for (int w = 0; w < 2; ++w) {
            TileQueue& Q = hq[pass * 2 + w];
            std::memset(&Q, 0, sizeof(Q));
            Q.seg = fb->lists[pass][w].seg;
            for (int x = 0; x < 8; ++x) Q.len[x] = fb->lists[pass][w].len[x];
        }
    fb->d_queues.ensure(sizeof(TileQueue) * hq.size());
    ILH_HIP(hipMemcpyAsync(fb->d_queues.p, hq.data(), sizeof(TileQueue) * hq.size(), hipMemcpyHostToDevice, st));
    ILH_HIP(hipStreamSynchronize(st));   // the host copies are released with the vectors' owner only
    return fb.release();
}

void fused_destroy(FusedBatch* fb) { delete fb; }

const std::vector<BatchProb>& fused_probs(const FusedBatch* fb) { return fb->probs; }

namespace {

// one pass with an explicit configuration
void launch_pass(FusedBatch* fb, int pass, int sched, int mult, int tpb, int amp_cap)
{
    ilh_ctx* ctx = fb->ctx;
    const int w = (pass == PASS_AMP || tpb == 256) ? 0 : 1;
    const TileList& L = fb->lists[pass][w];
    if (L.ntiles == 0) return;
    const BatchProb* pr = fb->d_probs.as<BatchProb>();
    const unsigned* tiles = fb->d_tiles.as<unsigned>() + L.off;
    TileQueue* q = fb->d_queues.as<TileQueue>() + pass * 2 + w;
    const unsigned grid = sched == SCHED_DISPATCH ? (unsigned)(8 * L.seg)
                                                  : (unsigned)(std::max(1, mult) * std::max(8, (ctx->num_cus + 7) / 8 * 8));
    const size_t lds = (size_t)tpb * V * sizeof(double2);
    hipStream_t st = ctx->stream;
    const int cap = std::min(amp_cap, (int)((160 * 1024 - lds - 64) / sizeof(double)));
    const size_t lds_rows =
        lds + (cap > 0 ? sizeof(double) * (size_t)std::min(cap, fb->max_amp) : (size_t)0);
    auto go = [&](auto kern, int threads, size_t bytes, auto... args) {
        set_lds_batch((const void*)kern, bytes);
        kern<<<grid, threads, bytes, st>>>(pr, tiles, q, args...);
    };
    const bool queue = sched == SCHED_QUEUE;
    switch (pass) {
    case PASS_AMP:
        if (queue) k_batch_amp<SCHED_QUEUE><<<grid, 256, 0, st>>>(pr, tiles, q);
        else k_batch_amp<SCHED_DISPATCH><<<grid, 256, 0, st>>>(pr, tiles, q);
        break;
    case PASS_ROWS:
        if (tpb == 256) {
            if (queue) go(k_batch_rows<256, SCHED_QUEUE>, 256, lds_rows, cap);
            else go(k_batch_rows<256, SCHED_DISPATCH>, 256, lds_rows, cap);
        } else {
            if (queue) go(k_batch_rows<512, SCHED_QUEUE>, 512, lds_rows, cap);
            else go(k_batch_rows<512, SCHED_DISPATCH>, 512, lds_rows, cap);
        }
        break;
    case PASS_COLS:
        if (tpb == 256) {
            if (queue) go(k_batch_cols<256, SCHED_QUEUE>, 256, lds);
            else go(k_batch_cols<256, SCHED_DISPATCH>, 256, lds);
        } else {
            if (queue) go(k_batch_cols<512, SCHED_QUEUE>, 512, lds);
            else go(k_batch_cols<512, SCHED_DISPATCH>, 512, lds);
        }
        break;
    default:
        if (tpb == 256) {
            if (queue) go(k_batch_inv<256, SCHED_QUEUE>, 256, lds);
            else go(k_batch_inv<256, SCHED_DISPATCH>, 256, lds);
        } else {
            if (queue) go(k_batch_inv<512, SCHED_QUEUE>, 512, lds);
            else go(k_batch_inv<512, SCHED_DISPATCH>, 512, lds);
        }
    }
    ILH_HIP(hipGetLastError());
}

}  // namespace

void fused_run(FusedBatch* fb, int first_pass)
{
    const FusedCfg& c = fb->cfg;
    for (int pass = first_pass; pass < NPASS; ++pass)
        launch_pass(fb, pass, c.sched[pass], c.mult[pass], c.tpb[pass], c.amp_cap);
}

// Times every candidate of each pass (one untimed launch, then the mean of
// three between HIP events on the context stream), in pass order so that
// each pass reads what the previous one wrote, and keeps the fastest.
void fused_tune(FusedBatch* fb)
{
    ilh_ctx* ctx = fb->ctx;
    hipStream_t st = ctx->stream;
    for (hipEvent_t& e : ctx->tev)
        if (!e) ILH_HIP(hipEventCreate(&e));
    struct Cand {
        int sched, mult, tpb, cap;
    };
    for (int pass = 0; pass < NPASS; ++pass) {
        std::vector<Cand> cands;
        const int ntpb = pass == PASS_AMP ? 1 : 2;
        for (int w = 0; w < ntpb; ++w)
            for (int cap : (pass == PASS_ROWS ? std::vector<int>{0, 4096} : std::vector<int>{4096})) {
                cands.push_back({SCHED_DISPATCH, 1, TPBS[w], cap});
                for (int mult : {1, 2, 4}) cands.push_back({SCHED_QUEUE, mult, TPBS[w], cap});
            }
        float best = 1e30f;
        Cand pick = cands[0];
        for (const Cand& k : cands) {
            launch_pass(fb, pass, k.sched, k.mult, k.tpb, k.cap);
            ILH_HIP(hipEventRecord(ctx->tev[0], st));
            for (int r = 0; r < 3; ++r) launch_pass(fb, pass, k.sched, k.mult, k.tpb, k.cap);
            ILH_HIP(hipEventRecord(ctx->tev[1], st));
            ILH_HIP(hipEventSynchronize(ctx->tev[1]));
            float ms = 0.f;
            ILH_HIP(hipEventElapsedTime(&ms, ctx->tev[0], ctx->tev[1]));
            ms /= 3.f;
            if (ms < best) {
                best = ms;
                pick = k;
            }
        }
        fb->cfg.sched[pass] = pick.sched;
        fb->cfg.mult[pass] = pick.mult;
        fb->cfg.tpb[pass] = pick.tpb;
        if (pass == PASS_ROWS) fb->cfg.amp_cap = pick.cap;
        fb->tuned_ms[pass] = best;
        // leave the pass's outputs as the kept configuration writes them
        launch_pass(fb, pass, pick.sched, pick.mult, pick.tpb, pick.cap);
    }
    ILH_HIP(hipStreamSynchronize(st));
}

void fused_config(const FusedBatch* fb, int* cfg, float* ms)
{
    for (int p = 0; p < NPASS; ++p) {
        if (cfg) {
            cfg[3 * p] = fb->cfg.sched[p];
            cfg[3 * p + 1] = fb->cfg.mult[p];
            cfg[3 * p + 2] = fb->cfg.tpb[p];
        }
        if (ms) ms[p] = fb->tuned_ms[p];
    }
    if (cfg) cfg[12] = fb->cfg.amp_cap;
}

}  // namespace ilh
