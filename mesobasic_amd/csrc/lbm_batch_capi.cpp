// lbm_batch_capi.cpp -- host side of the batched / serving form: job
// admission against an HBM budget, device job and work tables, ticks of
// steps (captured once in a hipGraph and replayed), retirement and the
// launch-configuration tuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "../../include/mesobasic_lbm.h"
#include "lbm_batch.h"
#include "lbm_internal.h"

using lbm::batch::Item;
using lbm::batch::Job;
using lbm::batch::Tables;

namespace {

#define BHIP(expr)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return lbm::set_error(LBM_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                                  __FILE__, __LINE__);                                       \
    } while (0)

// Kernel classes of one step: D2Q9 row-marching kernels by row width, the
// two-pass kernels (psi then collide) of each stencil and the D3Q19 march;
// kBlk9 the D2Q9 temporal blocks (several steps per launch); kMass is the
// per-tick mass check that follows the steps of a tick.
// Row classes by row width: a workgroup has as many threads as its class's
// width.  Finer classes (every multiple of 64 nodes, which leave fewer idle
// wavefronts per row) measured slower on MI355X with concurrent classes: 512
// D2Q9 jobs of 64..512 per side, row kernel, GLUPS 95-98 with these widths,
// 92-94 adding 384, 86-88 adding 192, 384 and 768.
constexpr int kNumRow = 5;
constexpr int kRowWidth[kNumRow] = {64, 128, 256, 512, 1024};
enum Cls { kRow0, kTp9Psi = kRow0 + kNumRow, kTp9Col, kTp19Psi, kTp19Col, kMarch19, kBlk9, kMass, kNumCls };
// Nodes per two-pass work item.  Serving stream (200 requests, 77% of the
// work in ragged D3Q19 jobs on the two passes), GLUPS / p99 s by item size:
// 1024 30.5 / 16.6, 2048 33.2 / 15.1, 4096 32.7 / 15.4, 8192 31.5 / 16.0,
// 16384 29.0 / 17.4, 32768 24.4 / 20.6: short items balance the last
// round of each persistent launch, very short ones pay per-item overhead.
constexpr int64_t kTpItem = 2048;
constexpr int kMarchChunk = 32;     // x planes per D3Q19 march item
constexpr int64_t kMassItem = 8192; // nodes per mass-check item
// D2Q9 row items: rounds of resident workgroups a row class's items should
// fill before its chunk is shortened, and the shortest chunk.  Serving
// stream (200 requests, D2Q9 128..1024 per side): p50 4.66 -> 4.50 s, p99
// 16.3 -> 15.9 s at 2 rounds; 512 jobs of 64..512 per side unchanged (66
// GLUPS); at 4 or 8 rounds the tuner measured the row kernel slower than
// the two passes on those jobs (61 GLUPS).
constexpr int kRowRounds = 2;
constexpr int kMinRowChunk = 2;

struct HostJob {
    int64_t id;
    lbm_job_desc d;
    int64_t n, ps;
    void* P[2] = {nullptr, nullptr};
    float* psi = nullptr;
    float* rho0 = nullptr;
    float* R[2] = {nullptr, nullptr};   // node shift: reference densities of P[0] / P[1]
    int64_t bytes = 0;
    int64_t remaining;
    int first = 1;
    bool finished = false;
    bool failed = false;        // mass went non-finite; retired at the tick that saw it
    int64_t done = 0;           // steps run
    int64_t steps_finite = 0;   // steps after which the last mass check was finite
    double mass = 0.0;          // FP64 total mass at the last check
    bool mass_valid = false;
    int waited = 0;             // ticks in a row it was resident and not advanced
};

int64_t plane_stride(int64_t n, int elem)
{
    const int64_t per = 256 / elem;
    int64_t ps = (n + per - 1) / per * per;
    if ((ps * elem) % 4096 == 0) ps += 512 / elem;
    return ps;
}

int elem_bytes(int prec) { return prec == LBM_FP32 ? 4 : 2; }

bool node_shift(int prec) { return prec == LBM_FP16N || prec == LBM_BF16N; }

}  // namespace

struct lbm_batch {
    lbm_batch_desc d{};
    hipStream_t stream = nullptr;
    int cus = 256;
    std::map<int64_t, std::unique_ptr<HostJob>> jobs;   // admitted, not collected
    std::vector<HostJob*> order;                        // device table order
    int64_t next_id = 1;
    int64_t in_use = 0;
    int64_t steps_run = 0;
    int64_t work_limit = 0;       // max nodes advanced per step (0: all)
    int max_wait = 0;             // ticks a job may wait under the work limit (0: no bound)
    // device tables
    Job* d_jobs = nullptr;
    int cap_jobs = 0;
    Item* d_items[kNumCls] = {};
    int cap_items[kNumCls] = {};
    int* d_nitems = nullptr;      // [kNumCls]
    unsigned* d_tickets = nullptr; // [kNumCls][2]
    int nitems[kNumCls] = {};
    unsigned class_mask = 0;      // classes with items; only these are launched
    int64_t class_nodes[kNumCls] = {};   // nodes of the jobs of each launch chain
    bool dirty = true;
    // launch configuration
    int chunk = 32;               // D2Q9 rows per row-kernel item
    int rows_kernel = 1;          // D2Q9 jobs on the row kernel (1), two-pass (0) or temporal blocks (2)
    int march19 = 1;              // D3Q19 kernel mode: 0 two-pass, 1 by shape, 2 march where it fits
    int tune_left = 0;            // ticks left to try candidate configurations
    struct Config {
        int rows_kernel, chunk;
    };
    std::vector<Config> candidates;
    std::vector<double> cand_rate;
    int cand_idx = 0;
    // a tick launched by lbm_batch_tick_begin and not yet ended
    bool tick_pending = false;
    std::chrono::steady_clock::time_point tick_t0;
    std::vector<int64_t> tick_done_before;
    // tick graph
    hipGraphExec_t graph = nullptr;
    bool graph_valid = false;
    // launch chains: the steps of each kernel class go to a stream of their
    // own (default priority), forked from `stream` at the start of a tick and
    // joined into it before the mass check.  A class's jobs depend only on
    // its own launches, so the classes of a mix overlap on the GPU: a class of
    // a few small jobs, whose launches are short serial chains on an almost
    // idle GPU, runs beside the large classes instead of after them.
    hipStream_t cstream[kNumCls] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[kNumCls] = {};
    bool concurrent = true;
    // pinned staging for per-tick counters
    Job* h_jobs = nullptr;
    // per-tick mass check: one FP64 partial per kMass item, the job (table
    // index) of each item, and the pinned copy the host adds up
    double* d_mass = nullptr;
    double* h_mass = nullptr;
    int cap_mass = 0;
    std::vector<int> mass_job;
};

namespace {

int row_class(int ny)
{
    for (int k = 0; k < kNumRow; k++)
        if (ny <= kRowWidth[k]) return k;
    return -1;
}

// D2Q9 temporal blocks: the constant-shift forms (the node-shift forms keep
// the row kernel).
bool use_blocks(const lbm_batch* b, const HostJob& j)
{
    return j.d.q == 9 && b->rows_kernel == 2 && !node_shift(b->d.precision);
}

bool use_rows(const lbm_batch* b, const HostJob& j)
{
    return j.d.q == 9 && b->rows_kernel && !use_blocks(b, j) && row_class(j.d.ny) >= 0;
}

// D3Q19 kernel of a job -- a function of its own shape and the batch's mode
// only, so a job runs the same kernel in any mix.  Mode 1 (default): the
// march for node-shift jobs whose y-z cross-section fills at least
// kMarchMinFill of the march tiles it needs, the two passes for the others
// and for the constant-shift forms.  The march's partial tiles cost in
// proportion to the unfilled part.  Measured on MI355X (FP16 node shift,
// GLUPS, march vs two passes): 128^3 jobs 35.7 vs 27.4; the 64 D3Q19
// requests of the serving stream (64..192 per side) 21.9 vs 27.5; constant
// shift, the same requests 23.8 vs 34.9.  2: the march wherever it fits;
// 0: the two passes.
constexpr double kMarchMinFill = 0.9;

bool use_march(const lbm_batch* b, const HostJob& j)
{
    const bool fits = j.d.q == 19 && j.d.nx >= 2 && j.d.ny >= lbm::batch::kMarchTY &&
                      j.d.nz >= lbm::batch::kMarchTZ;
    if (!fits || b->march19 == 0) return false;
    if (b->march19 == 2) return true;
    if (!node_shift(b->d.precision)) return false;
    const int64_t ty = lbm::batch::kMarchTY, tz = lbm::batch::kMarchTZ;
    const double cover = static_cast<double>((j.d.ny + ty - 1) / ty * ty) *
                         static_cast<double>((j.d.nz + tz - 1) / tz * tz);
    return static_cast<double>(j.d.ny) * j.d.nz >= kMarchMinFill * cover;
}

// Launch chain of a job: the kernel class that steps it (the two-pass
// classes by their psi pass).
int chain_of(const lbm_batch* b, const HostJob& j)
{
    if (use_blocks(b, j)) return kBlk9;
    if (use_rows(b, j)) return row_class(j.d.ny);
    if (use_march(b, j)) return kMarch19;
    return j.d.q == 9 ? kTp9Psi : kTp19Psi;
}

int ensure_items(lbm_batch* b, int cls, int n)
{
    if (b->cap_items[cls] >= n) return LBM_OK;
    const int cap = std::max(n, 2 * b->cap_items[cls] + 64);
    if (b->d_items[cls]) (void)hipFree(b->d_items[cls]);
    b->d_items[cls] = nullptr;
    b->cap_items[cls] = 0;
    BHIP(lbm::dev_malloc(&b->d_items[cls], sizeof(Item) * cap));
    b->cap_items[cls] = cap;
    b->graph_valid = false;   // kernel arguments hold the table pointers
    return LBM_OK;
}

int ensure_jobs(lbm_batch* b, int n)
{
    if (b->cap_jobs >= n) return LBM_OK;
    const int cap = std::max(n, 2 * b->cap_jobs + 64);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    b->d_jobs = nullptr;
    b->h_jobs = nullptr;
    b->cap_jobs = 0;
    BHIP(lbm::dev_malloc(&b->d_jobs, sizeof(Job) * cap));
    BHIP(hipHostMalloc(&b->h_jobs, sizeof(Job) * cap, hipHostMallocDefault));
    static_assert(sizeof(Job) % 4 == 0, "job table is cleared as 32-bit words");
    // unused table entries are cleared: the per-step counter kernel walks the
    // whole capacity
    BHIP(lbm::batch::zero_u32(reinterpret_cast<unsigned*>(b->d_jobs),
                              static_cast<int>(sizeof(Job) * cap / 4), b->stream));
    b->cap_jobs = cap;
    b->graph_valid = false;
    return LBM_OK;
}

int ensure_mass(lbm_batch* b, int n)
{
    if (b->cap_mass >= n) return LBM_OK;
    const int cap = std::max(n, 2 * b->cap_mass + 256);
    if (b->d_mass) (void)hipFree(b->d_mass);
    if (b->h_mass) (void)hipHostFree(b->h_mass);
    b->d_mass = nullptr;
    b->h_mass = nullptr;
    b->cap_mass = 0;
    BHIP(lbm::dev_malloc(&b->d_mass, sizeof(double) * cap));
    BHIP(hipHostMalloc(&b->h_mass, sizeof(double) * cap, hipHostMallocDefault));
    b->cap_mass = cap;
    b->graph_valid = false;   // the mass kernel's argument holds the pointer
    return LBM_OK;
}

Job device_view(const HostJob& j, int prec, double rs, bool blk = false)
{
    Job v{};
    v.blk = blk ? 1 : 0;
    v.P[0] = j.P[0];
    v.P[1] = j.P[1];
    v.psi = j.psi;
    v.R[0] = j.R[0];
    v.R[1] = j.R[1];
    v.rho0 = (j.d.q == 9) ? j.rho0 : nullptr;   // LBfluid2D-order first step
    v.nx = j.d.nx;
    v.ny = j.d.ny;
    v.nz = j.d.q == 9 ? 1 : j.d.nz;
    // jobs join (and resume after lbm_batch_extend) at tick boundaries, where
    // the launch parity is 0: that launch must read the current buffer
    v.phase = static_cast<int>(j.done & 1);
    v.n = j.n;
    v.ps = j.ps;
    v.tau = static_cast<float>(j.d.tau);
    v.G = static_cast<float>(j.d.G);
    v.rs = static_cast<float>((prec == LBM_FP32 || node_shift(prec)) ? 0.0 : rs);
    v.first = j.first;
    v.remaining = j.remaining;
    v.done = j.done;
    v.q = j.d.q;
    return v;
}

int persistent_grid(const lbm_batch* b, int threads)
{
    // enough workgroups to fill every CU at the kernel's occupancy
    const int per_cu = std::max(1, 2048 / threads);
    return b->cus * std::min(per_cu, 8);
}

// x rows per item of each row class (rows[k]: the rows of class k): the
// tuned chunk, shortened (to kMinRowChunk at least) until the items fill
// kRowRounds rounds of resident workgroups.  A few narrow jobs otherwise run
// as a few long items -- each a serial chain of row-load round trips -- on
// an almost idle GPU.  Classes that run one after another (no concurrency)
// each fill their own rounds; concurrent classes share the GPU, so their
// items together fill rounds of its resident threads (2048 per CU) and a
// class is not cut into short items, each of which re-reads three rows,
// only because it is small.  Results do not depend on the chunk.
void row_chunks(const lbm_batch* b, const int64_t (&rows)[kNumRow], int (&chunk)[kNumRow])
{
    const int lo = std::min(kMinRowChunk, b->chunk);
    if (b->concurrent) {
        int64_t threads = 0;   // row threads of all items of one row each
        for (int k = 0; k < kNumRow; k++) threads += rows[k] * kRowWidth[k];
        const int64_t c = threads / std::max<int64_t>(1, int64_t(kRowRounds) * b->cus * 2048);
        for (int k = 0; k < kNumRow; k++) chunk[k] = static_cast<int>(std::clamp<int64_t>(c, lo, b->chunk));
        return;
    }
    for (int k = 0; k < kNumRow; k++) {
        const int64_t per_round = static_cast<int64_t>(kRowRounds) * persistent_grid(b, kRowWidth[k]);
        const int64_t c = rows[k] / std::max<int64_t>(1, per_round);
        chunk[k] = static_cast<int>(std::clamp<int64_t>(c, lo, b->chunk));
    }
}

// Error of a call that would change a batch while its tick is in flight.
int in_flight(const char* what)
{
    return lbm::set_error(LBM_ESTATE, "%s: a tick is in flight (lbm_batch_tick_end first)", what);
}

// Rebuilds and uploads the job and item tables (between ticks only).
int upload_tables(lbm_batch* b)
{
    b->order.clear();
    for (auto& kv : b->jobs)
        if (!kv.second->finished) b->order.push_back(kv.second.get());
    if (b->work_limit > 0 && !b->order.empty()) {
        // shortest remaining work first, within the node limit; jobs that
        // have waited max_wait ticks go first, longest wait first.  A running
        // job's remaining work only shrinks, so the set changes only when a
        // job is submitted, finishes or reaches the wait bound (each marks
        // the tables dirty)
        const int mw = b->max_wait;
        std::stable_sort(b->order.begin(), b->order.end(), [mw](const HostJob* a, const HostJob* z) {
            const bool sa = mw > 0 && a->waited >= mw, sz = mw > 0 && z->waited >= mw;
            if (sa != sz) return sa;
            if (sa) return a->waited > z->waited;
            return a->n * a->remaining < z->n * z->remaining;
        });
        int64_t nodes = 0;
        size_t keep = 0;
        while (keep < b->order.size() && (keep == 0 || nodes + b->order[keep]->n <= b->work_limit))
            nodes += b->order[keep++]->n;
        b->order.resize(keep);
    }
    int rc = ensure_jobs(b, static_cast<int>(b->order.size()) + 1);
    if (rc) return rc;
    std::vector<Item> items[kNumCls];
    b->mass_job.clear();
    int chunk[kNumRow];
    {
        int64_t rows[kNumRow] = {};
        for (const HostJob* j : b->order)
            if (use_rows(b, *j)) rows[row_class(j->d.ny)] += j->d.nx;
        row_chunks(b, rows, chunk);
    }
    for (int ji = 0; ji < static_cast<int>(b->order.size()); ji++) {
        const HostJob& j = *b->order[ji];
        b->h_jobs[ji] = device_view(j, b->d.precision, b->d.rho_shift, use_blocks(b, j));
        for (int64_t c0 = 0; c0 < j.n; c0 += kMassItem) {
            Item it{};
            it.job = ji;
            it.c0 = c0;
            it.c1 = std::min(j.n, c0 + kMassItem);
            items[kMass].push_back(it);
            b->mass_job.push_back(ji);
        }
        if (use_blocks(b, j)) {
            for (int x0 = 0; x0 < j.d.nx; x0 += lbm::batch::kBlkT)
                for (int y0 = 0; y0 < j.d.ny; y0 += lbm::batch::kBlkT) {
                    Item it{};
                    it.job = ji;
                    it.x0 = x0;
                    it.y0 = y0;
                    items[kBlk9].push_back(it);
                }
        } else if (use_rows(b, j)) {
            const int cls = row_class(j.d.ny);
            for (int x0 = 0; x0 < j.d.nx; x0 += chunk[cls]) {
                Item it{};
                it.job = ji;
                it.x0 = x0;
                it.x1 = std::min(j.d.nx, x0 + chunk[cls]);
                items[cls].push_back(it);
            }
        } else if (use_march(b, j)) {
            for (int y0 = 0; y0 < j.d.ny; y0 += lbm::batch::kMarchTY)
                for (int z0 = 0; z0 < j.d.nz; z0 += lbm::batch::kMarchTZ)
                    for (int x0 = 0; x0 < j.d.nx; x0 += kMarchChunk) {
                        Item it{};
                        it.job = ji;
                        it.y0 = y0;
                        it.z0 = z0;
                        it.x0 = x0;
                        it.x1 = std::min(j.d.nx, x0 + kMarchChunk);
                        items[kMarch19].push_back(it);
                    }
        } else {
            const int cp = j.d.q == 9 ? kTp9Psi : kTp19Psi;
            for (int64_t c0 = 0; c0 < j.n; c0 += kTpItem) {
                Item it{};
                it.job = ji;
                it.c0 = c0;
                it.c1 = std::min(j.n, c0 + kTpItem);
                items[cp].push_back(it);
                items[cp + 1].push_back(it);
            }
        }
    }
    rc = ensure_mass(b, static_cast<int>(items[kMass].size()) + 1);
    if (rc) return rc;
    // largest jobs first, so that the long items start early (the mass items
    // stay in job order: the host adds each job's partials in that order)
    unsigned mask = 0;
    for (int c = 0; c < kNumCls; c++) {
        if (c != kMass)
            std::stable_sort(items[c].begin(), items[c].end(), [&](const Item& a, const Item& z) {
                return b->order[a.job]->n > b->order[z.job]->n;
            });
        rc = ensure_items(b, c, static_cast<int>(items[c].size()) + 1);
        if (rc) return rc;
        b->nitems[c] = static_cast<int>(items[c].size());
        if (!items[c].empty()) mask |= 1u << c;
        if (!items[c].empty())
            BHIP(hipMemcpyAsync(b->d_items[c], items[c].data(), sizeof(Item) * items[c].size(),
                                hipMemcpyHostToDevice, b->stream));
    }
    if (!b->order.empty())
        BHIP(hipMemcpyAsync(b->d_jobs, b->h_jobs, sizeof(Job) * b->order.size(),
                            hipMemcpyHostToDevice, b->stream));
    BHIP(hipMemcpyAsync(b->d_nitems, b->nitems, sizeof(int) * kNumCls, hipMemcpyHostToDevice,
                        b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    if (mask != b->class_mask) b->graph_valid = false;   // the launch sequence changes
    b->class_mask = mask;
    std::fill(std::begin(b->class_nodes), std::end(b->class_nodes), int64_t(0));
    for (const HostJob* j : b->order) b->class_nodes[chain_of(b, *j)] += j->n;
    b->dirty = false;
    return LBM_OK;
}

int launch_step(lbm_batch* b, int step);
int launch_mass(lbm_batch* b);

// Temporal-block launches of a tick: an even number (the ticket parity of
// the class returns to 0 for the next tick) of at least steps_per_tick /
// kBlkK, so block jobs advance at least steps_per_tick steps per tick.
int block_launches(const lbm_batch* b)
{
    const int k = lbm::batch::kBlkK;
    const int l = (b->d.steps_per_tick + k - 1) / k;
    return l + (l & 1);
}

int launch_blocks(lbm_batch* b, int parity, hipStream_t s)
{
    Tables t{b->d_jobs, b->d_items[kBlk9], b->d_nitems + kBlk9, b->d_tickets + 2 * kBlk9};
    BHIP(lbm::batch::d2q9_block_step(b->d.precision, t, parity, persistent_grid(b, 1024), s));
    BHIP(lbm::batch::advance_blocks(b->d_jobs, b->cap_jobs, b->d_tickets + 2 * kMass, s));
    return LBM_OK;
}

// The launches of step `step` of the tick for chain c (a row class, a
// two-pass class by its psi pass, or the D3Q19 march) on stream s.
int launch_class(lbm_batch* b, int c, int step, hipStream_t s)
{
    if (c < kNumRow) {
        Tables t{b->d_jobs, b->d_items[c], b->d_nitems + c, b->d_tickets + 2 * c};
        BHIP(lbm::batch::d2q9_rows_step(b->d.precision, t, step, kRowWidth[c],
                                        persistent_grid(b, kRowWidth[c]), s));
    } else if (c == kTp9Psi || c == kTp19Psi) {
        const int q = c == kTp9Psi ? 9 : 19;
        Tables tp{b->d_jobs, b->d_items[c], b->d_nitems + c, b->d_tickets + 2 * c};
        Tables tc{b->d_jobs, b->d_items[c + 1], b->d_nitems + c + 1, b->d_tickets + 2 * (c + 1)};
        BHIP(lbm::batch::twopass_psi(b->d.precision, q, tp, step, persistent_grid(b, 256), s));
        BHIP(lbm::batch::twopass_collide(b->d.precision, q, tc, step, persistent_grid(b, 256), s));
    } else if (c == kMarch19) {
        Tables tm{b->d_jobs, b->d_items[kMarch19], b->d_nitems + kMarch19, b->d_tickets + 2 * kMarch19};
        BHIP(lbm::batch::d3q19_march_step(b->d.precision, tm, step,
                                          persistent_grid(b, lbm::batch::kMarchTY * lbm::batch::kMarchTZ), s));
    }
    return LBM_OK;
}

// Chains of the current tables, the largest (by nodes) first.
std::vector<int> tick_chains(const lbm_batch* b)
{
    std::vector<int> c;
    for (int k = kRow0; k < kRow0 + kNumRow; k++)
        if (b->class_mask & (1u << k)) c.push_back(k);
    for (int k : {kTp9Psi, kTp19Psi, kMarch19, kBlk9})
        if (b->class_mask & (1u << k)) c.push_back(k);
    std::stable_sort(c.begin(), c.end(),
                     [b](int x, int y) { return b->class_nodes[x] > b->class_nodes[y]; });
    return c;
}

// Launch units of chain c in one tick: its block launches (each followed by
// its counter update), or the class's kernels of each step of the tick.
int chain_units(const lbm_batch* b, int c)
{
    return c == kBlk9 ? block_launches(b) : b->d.steps_per_tick;
}

// Unit m of chain c on stream s.
int launch_unit(lbm_batch* b, int c, int m, hipStream_t s)
{
    if (c == kBlk9) return launch_blocks(b, m & 1, s);
    return launch_class(b, c, m, s);
}

// The launches of a tick: each chain's units (on its own stream when there
// are several and chains run concurrently), the counter update of the
// single-step jobs, then the mass check, on the batch's stream.  Concurrent chains are issued unit by unit across the
// chains, the largest chain first: a hipGraph replay submits its kernel
// nodes in about the order they were captured, and captured chain after
// chain the largest class started only after the small ones had run most of
// their steps alone on an almost idle GPU (MI355X kernel trace, 512 D2Q9
// jobs: the 64-wide class ran 13 steps alone before the next chain began).
int launch_tick(lbm_batch* b)
{
    const std::vector<int> chains = tick_chains(b);
    if (chains.size() > 1 && b->concurrent) {
        BHIP(hipEventRecord(b->ev_fork, b->stream));
        int units = 0;
        for (int c : chains) {
            BHIP(hipStreamWaitEvent(b->cstream[c], b->ev_fork, 0));
            units = std::max(units, chain_units(b, c));
        }
        int rc = LBM_OK;
        for (int m = 0; m < units && rc == LBM_OK; m++)
            for (int c : chains)
                if (rc == LBM_OK && m < chain_units(b, c)) rc = launch_unit(b, c, m, b->cstream[c]);
        // every forked stream is joined back, also after a failed launch (a
        // capture cannot end with a stream left out)
        for (int c : chains) {
            BHIP(hipEventRecord(b->ev_join[c], b->cstream[c]));
            BHIP(hipStreamWaitEvent(b->stream, b->ev_join[c], 0));
        }
        if (rc) return rc;
    } else {
        if (b->class_mask & (1u << kBlk9))
            for (int m = 0; m < chain_units(b, kBlk9); m++) {
                int rc = launch_unit(b, kBlk9, m, b->stream);
                if (rc) return rc;
            }
        if (std::any_of(chains.begin(), chains.end(), [](int c) { return c != kBlk9; }))
            for (int st = 0; st < b->d.steps_per_tick; st++) {
                int rc = launch_step(b, st);
                if (rc) return rc;
            }
    }
    if (std::any_of(chains.begin(), chains.end(), [](int c) { return c != kBlk9; }))
        BHIP(lbm::batch::advance_counters(b->d_jobs, b->cap_jobs, b->d.steps_per_tick,
                                          b->d_tickets + 2 * kMass, b->stream));
    return launch_mass(b);
}

// Step `step` of the tick for every single-step job, all classes on the
// batch's stream.
int launch_step(lbm_batch* b, int step)
{
    for (int c : tick_chains(b)) {
        if (c == kBlk9) continue;
        int rc = launch_class(b, c, step, b->stream);
        if (rc) return rc;
    }
    return LBM_OK;
}

// The tick's mass check, after its last step (the last advance_counters
// launch has cleared its ticket counter).
int launch_mass(lbm_batch* b)
{
    if (!(b->class_mask & (1u << kMass))) return LBM_OK;
    Tables tm{b->d_jobs, b->d_items[kMass], b->d_nitems + kMass, nullptr};
    BHIP(lbm::batch::job_mass(b->d.precision, tm, b->d_tickets + 2 * kMass, b->d_mass,
                              persistent_grid(b, 256), b->stream));
    return LBM_OK;
}

// Captures the K steps of a tick into a graph (when graphs are on and the
// launch sequence changed since the last capture).
int prepare_tick(lbm_batch* b)
{
    const int K = b->d.steps_per_tick;
    if (b->d.use_graph) {
        if (!b->graph_valid) {
            if (b->graph) (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
            hipGraph_t g = nullptr;
            BHIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = launch_tick(b);
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (e != hipSuccess) return lbm::set_error(LBM_EHIP, "graph capture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return lbm::set_error(LBM_EHIP, "graph instantiate: %s", hipGetErrorString(e));
            b->graph_valid = true;
        }
    }
    return LBM_OK;
}

int run_tick(lbm_batch* b)
{
    if (b->d.use_graph) {
        BHIP(hipGraphLaunch(b->graph, b->stream));
        return LBM_OK;
    }
    return launch_tick(b);
}

void free_job(lbm_batch* b, HostJob* j)
{
    for (void* p : j->P)
        if (p) (void)hipFree(p);
    if (j->psi) (void)hipFree(j->psi);
    if (j->rho0) (void)hipFree(j->rho0);
    for (float* r : j->R)
        if (r) (void)hipFree(r);
    b->in_use -= j->bytes;
}

int64_t job_bytes(const lbm_batch* b, const lbm_job_desc& d)
{
    const int64_t n = int64_t(d.nx) * d.ny * (d.q == 9 ? 1 : d.nz);
    const int elem = elem_bytes(b->d.precision);
    // populations, psi, rho0, and the node-shift reference densities
    return 2 * plane_stride(n, elem) * d.q * elem + 8 * n + (node_shift(b->d.precision) ? 8 * n : 0);
}

}  // namespace

extern "C" {

int lbm_batch_create(const lbm_batch_desc* desc, lbm_batch** out)
{
    if (!desc || !out) return lbm::set_error(LBM_EINVAL, "lbm_batch_create: NULL argument");
    *out = nullptr;
    if (desc->precision != LBM_FP32 && desc->precision != LBM_FP16S && desc->precision != LBM_BF16S &&
        !node_shift(desc->precision))
        return lbm::set_error(LBM_EINVAL,
                              "batch precision must be LBM_FP32, LBM_FP16S, LBM_BF16S, LBM_FP16N or LBM_BF16N");
    if (desc->steps_per_tick < 1) return lbm::set_error(LBM_EINVAL, "steps_per_tick must be >= 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return lbm::set_error(LBM_ENODEV, "lbm_batch_create: no HIP device");
    auto b = std::make_unique<lbm_batch>();
    b->d = *desc;
    b->d.steps_per_tick += b->d.steps_per_tick & 1;   // even: every tick starts at parity 0
    int dev = 0;
    BHIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    BHIP(hipGetDeviceProperties(&prop, dev));
    b->cus = prop.multiProcessorCount;
    if (b->d.hbm_budget <= 0) {
        size_t free_b = 0, total_b = 0;
        BHIP(hipMemGetInfo(&free_b, &total_b));
        b->d.hbm_budget = static_cast<int64_t>(0.9 * static_cast<double>(free_b));
    }
    BHIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (int c = 0; c < kNumCls; c++) {
        BHIP(hipStreamCreateWithFlags(&b->cstream[c], hipStreamNonBlocking));
        BHIP(hipEventCreateWithFlags(&b->ev_join[c], hipEventDisableTiming));
    }
    BHIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    BHIP(lbm::dev_malloc(&b->d_nitems, sizeof(int) * kNumCls));
    BHIP(lbm::dev_malloc(&b->d_tickets, sizeof(unsigned) * 2 * kNumCls));
    BHIP(lbm::batch::zero_u32(b->d_tickets, 2 * kNumCls, b->stream));
    BHIP(hipStreamSynchronize(b->stream));
    if (b->d.autotune) {
        // D2Q9 row kernel with row chunks of 16 / 32 / 64, or two-pass (the
        // D3Q19 kernel is not tuned: march and two-pass differ in rounding)
        for (int rk : {1, 0})
            for (int ch : {16, 32, 64})
                if (rk || ch == 32) b->candidates.push_back({rk, ch});
        if (!node_shift(b->d.precision)) b->candidates.push_back({2, 32});   // temporal blocks
        b->cand_rate.assign(b->candidates.size(), 0.0);
        b->tune_left = static_cast<int>(b->candidates.size());
        b->rows_kernel = b->candidates[0].rows_kernel;
        b->chunk = b->candidates[0].chunk;
    }
    *out = b.release();
    return LBM_OK;
}

void lbm_batch_destroy(lbm_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& kv : b->jobs) free_job(b, kv.second.get());
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    for (Item* p : b->d_items)
        if (p) (void)hipFree(p);
    if (b->d_nitems) (void)hipFree(b->d_nitems);
    if (b->d_tickets) (void)hipFree(b->d_tickets);
    if (b->d_mass) (void)hipFree(b->d_mass);
    if (b->h_mass) (void)hipHostFree(b->h_mass);
    for (int c = 0; c < kNumCls; c++) {
        if (b->cstream[c]) (void)hipStreamSynchronize(b->cstream[c]);
        if (b->cstream[c]) (void)hipStreamDestroy(b->cstream[c]);
        if (b->ev_join[c]) (void)hipEventDestroy(b->ev_join[c]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int64_t lbm_batch_job_bytes(const lbm_batch* b, const lbm_job_desc* job)
{
    return (b && job) ? job_bytes(b, *job) : -1;
}

int lbm_batch_submit(lbm_batch* b, const lbm_job_desc* job, const double* rho0, int64_t* job_id)
{
    if (!b || !job || !rho0) return lbm::set_error(LBM_EINVAL, "lbm_batch_submit: NULL argument");
    if (b->tick_pending) return in_flight("lbm_batch_submit");
    if (job->q != 9 && job->q != 19) return lbm::set_error(LBM_EINVAL, "job q must be 9 or 19");
    const int nz = job->q == 9 ? 1 : job->nz;
    if (job->nx < 3 || job->ny < 3 || nz < 1 || (job->q == 19 && nz < 3))
        return lbm::set_error(LBM_EINVAL, "job lattice too small (%d x %d x %d)", job->nx, job->ny, nz);
    if (!(job->tau > 0.5) || !std::isfinite(job->G) || job->steps < 0)
        return lbm::set_error(LBM_EINVAL, "job needs tau > 0.5, finite G, steps >= 0");
    const int64_t n = int64_t(job->nx) * job->ny * nz;
    // the batch kernels address a job's arrays with 32-bit byte offsets of
    // 4-byte values
    if (n >= (int64_t(1) << 30))
        return lbm::set_error(LBM_EINVAL, "job too large (%lld nodes; a batch job holds < 2^30)",
                              static_cast<long long>(n));
    const int64_t bytes = job_bytes(b, *job);
    if (b->in_use + bytes > b->d.hbm_budget)
        return lbm::set_error(LBM_ENOMEM, "job needs %lld B, %lld B of the %lld B budget left",
                              (long long)bytes, (long long)(b->d.hbm_budget - b->in_use),
                              (long long)b->d.hbm_budget);
    auto j = std::make_unique<HostJob>();
    j->id = b->next_id++;
    j->d = *job;
    j->d.nz = nz;
    j->n = n;
    j->ps = plane_stride(n, elem_bytes(b->d.precision));
    j->remaining = job->steps;
    j->bytes = bytes;
    const int64_t pop = j->ps * job->q * elem_bytes(b->d.precision);
    auto undo = [&](int code) {
        free_job(b, j.get());
        b->in_use += j->bytes;   // free_job subtracted what was never added
        return code;
    };
    for (int k = 0; k < 2; k++)
        if (lbm::dev_malloc(&j->P[k], pop) != hipSuccess)
            return undo(lbm::set_error(LBM_ENOMEM, "hipMalloc of job populations failed"));
    if (lbm::dev_malloc(&j->psi, n * 4) != hipSuccess || lbm::dev_malloc(&j->rho0, n * 4) != hipSuccess)
        return undo(lbm::set_error(LBM_ENOMEM, "hipMalloc of job fields failed"));
    if (node_shift(b->d.precision))
        for (int k = 0; k < 2; k++)
            if (lbm::dev_malloc(&j->R[k], n * 4) != hipSuccess)
                return undo(lbm::set_error(LBM_ENOMEM, "hipMalloc of job reference density failed"));
    std::vector<float> r32(static_cast<size_t>(n));
    for (int64_t c = 0; c < n; c++) r32[c] = static_cast<float>(rho0[c]);
    if (hipMemcpyAsync(j->rho0, r32.data(), n * 4, hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return undo(lbm::set_error(LBM_EHIP, "upload of rho0 failed"));
    Job v = device_view(*j, b->d.precision, b->d.rho_shift);
    v.rho0 = j->rho0;   // the initialisation reads it for both stencils
    if (lbm::batch::init_job(b->d.precision, job->q, v, 0, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return undo(lbm::set_error(LBM_EHIP, "job initialisation failed"));
    b->in_use += bytes;
    *job_id = j->id;
    if (job->steps == 0) j->finished = true;
    b->jobs[j->id] = std::move(j);
    b->dirty = true;
    return LBM_OK;
}

int lbm_batch_tick_begin(lbm_batch* b)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_tick_begin: NULL batch");
    if (b->tick_pending) return lbm::set_error(LBM_ESTATE, "lbm_batch_tick_begin: a tick is already in flight");
    if (b->dirty) {
        int rc = upload_tables(b);
        if (rc) return rc;
    }
    if (!b->order.empty()) {
        int rc = prepare_tick(b);
        if (rc) return rc;
    }
    b->tick_t0 = std::chrono::steady_clock::now();
    b->tick_done_before.resize(b->order.size());
    for (size_t ji = 0; ji < b->order.size(); ji++) b->tick_done_before[ji] = b->order[ji]->done;
    if (!b->order.empty()) {
        int rc = run_tick(b);
        if (rc) return rc;
        BHIP(hipMemcpyAsync(b->h_jobs, b->d_jobs, sizeof(Job) * b->order.size(),
                            hipMemcpyDeviceToHost, b->stream));
        if (b->nitems[kMass] > 0)
            BHIP(hipMemcpyAsync(b->h_mass, b->d_mass, sizeof(double) * b->nitems[kMass],
                                hipMemcpyDeviceToHost, b->stream));
    }
    b->tick_pending = true;
    return LBM_OK;
}

int lbm_batch_tick(lbm_batch* b, int* n_finished)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_tick: NULL batch");
    int rc = lbm_batch_tick_begin(b);
    if (rc) return rc;
    return lbm_batch_tick_end(b, n_finished);
}

int lbm_batch_tick_end(lbm_batch* b, int* n_finished)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_tick_end: NULL batch");
    if (!b->tick_pending) return lbm::set_error(LBM_ESTATE, "lbm_batch_tick_end: no tick in flight");
    b->tick_pending = false;
    BHIP(hipStreamSynchronize(b->stream));
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - b->tick_t0).count();
    const std::vector<int64_t>& done_before = b->tick_done_before;
    int64_t nodes = 0;   // lattice updates of the tick
    for (size_t ji = 0; ji < b->order.size(); ji++)
        nodes += b->order[ji]->n * (b->h_jobs[ji].done - done_before[ji]);
    b->steps_run += b->d.steps_per_tick;
    std::vector<double> mass(b->order.size(), 0.0);
    for (int k = 0; k < b->nitems[kMass]; k++) mass[b->mass_job[k]] += b->h_mass[k];
    // waiting times of the resident jobs this tick did not advance
    for (auto& kv : b->jobs) kv.second->waited += kv.second->finished ? 0 : 1;
    for (HostJob* j : b->order) j->waited = 0;
    if (b->max_wait > 0)
        for (auto& kv : b->jobs)
            if (!kv.second->finished && kv.second->waited == b->max_wait) b->dirty = true;
    for (int ji = 0; ji < static_cast<int>(b->order.size()); ji++) {
        HostJob* j = b->order[ji];
        j->remaining = b->h_jobs[ji].remaining;
        j->first = b->h_jobs[ji].first;
        j->done = b->h_jobs[ji].done;
        j->mass = mass[ji];
        j->mass_valid = true;
        if (!std::isfinite(mass[ji])) {
            // a non-finite value in the job's populations: retire it now, so
            // it stops taking steps, and report it as failed
            j->failed = true;
            if (!j->finished) b->dirty = true;
            j->finished = true;
        } else {
            j->steps_finite = j->done;
        }
        if (j->remaining == 0 && !j->finished) {
            j->finished = true;
            b->dirty = true;   // finished jobs leave the work tables
        }
    }
    // tuner: one candidate configuration per tick, then the fastest
    if (b->tune_left > 0 && nodes > 0) {
        b->cand_rate[b->cand_idx] = nodes / std::max(secs, 1e-9);
        b->tune_left--;
        b->cand_idx++;
        const size_t pick = b->tune_left > 0
                                ? static_cast<size_t>(b->cand_idx)
                                : static_cast<size_t>(std::max_element(b->cand_rate.begin(),
                                                                       b->cand_rate.end()) -
                                                      b->cand_rate.begin());
        b->rows_kernel = b->candidates[pick].rows_kernel;
        b->chunk = b->candidates[pick].chunk;
        b->dirty = true;
        b->graph_valid = false;
    }
    int nf = 0;
    for (auto& kv : b->jobs) nf += kv.second->finished ? 1 : 0;
    if (n_finished) *n_finished = nf;
    return LBM_OK;
}

int lbm_batch_finished(lbm_batch* b, int64_t* ids, int max, int* n)
{
    if (!b || !n) return lbm::set_error(LBM_EINVAL, "lbm_batch_finished: NULL argument");
    int k = 0;
    for (auto& kv : b->jobs)
        if (kv.second->finished && k < max && ids) ids[k++] = kv.first;
    *n = k;
    return LBM_OK;
}

int lbm_batch_collect(lbm_batch* b, int64_t job_id, double* rho, double* u, double* v, double* w)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_collect: NULL batch");
    if (b->tick_pending) return in_flight("lbm_batch_collect");
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) return lbm::set_error(LBM_EINVAL, "unknown job %lld", (long long)job_id);
    HostJob* j = it->second.get();
    if (!j->finished) return lbm::set_error(LBM_ESTATE, "job %lld has not finished", (long long)job_id);
    const int64_t n = j->n;
    double* d = nullptr;
    BHIP(lbm::dev_malloc(&d, 4 * n * 8));
    Job view = device_view(*j, b->d.precision, b->d.rho_shift);
    const int buf = static_cast<int>(j->done & 1);   // the buffer the last step wrote
    hipError_t e = lbm::batch::job_moments(b->d.precision, j->d.q, view, buf, d, d + n, d + 2 * n,
                                           d + 3 * n, b->stream);
    double* outs[4] = {rho, u, v, w};
    for (int k = 0; k < 4 && e == hipSuccess; k++)
        if (outs[k]) e = hipMemcpyAsync(outs[k], d + k * n, n * 8, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return lbm::set_error(LBM_EHIP, "collect: %s", hipGetErrorString(e));
    const bool failed = j->failed;
    const long long id = static_cast<long long>(j->id), at = static_cast<long long>(j->done),
                    ok = static_cast<long long>(j->steps_finite);
    free_job(b, j);
    b->jobs.erase(it);
    if (failed)
        return lbm::set_error(LBM_EJOB,
                              "job %lld failed: non-finite mass after step %lld (finite after step %lld)",
                              id, at, ok);
    return LBM_OK;
}

int lbm_batch_read(lbm_batch* b, int64_t job_id, double* rho, double* u, double* v, double* w, double* fx,
                   double* fy, double* fz)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_read: NULL batch");
    if (b->tick_pending) return in_flight("lbm_batch_read");
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) return lbm::set_error(LBM_EINVAL, "unknown job %lld", (long long)job_id);
    const HostJob& j = *it->second;
    const int64_t n = j.n;
    const int64_t done = j.done;
    const bool order3 = (j.d.q == 19);   // D3Q19 jobs: scmp3D order; D2Q9: LBfluid2D order
    // device arrays: the reference's r,u,v,w; the density the last step
    // started from (its moments' u,v,w unused); psi; fx,fy,fz
    double* d = nullptr;
    BHIP(lbm::dev_malloc(&d, 12 * n * 8));
    double *R = d, *S = d + 4 * n, *psi = d + 8 * n, *F = d + 9 * n;
    Job view = device_view(j, b->d.precision, b->d.rho_shift);
    const int cur = static_cast<int>(done & 1), prev = 1 - cur;
    hipError_t e = hipSuccess;
    if (done == 0) {
        // rho0 and zero velocity before any step (scmp3D.cpp:63-76,
        // LBfluid2D::setRho / setU)
        e = lbm::batch::widen(j.rho0, R, n, b->stream);
        if (e == hipSuccess) e = hipMemsetAsync(R + n, 0, 3 * n * 8, b->stream);
    } else {
        e = lbm::batch::job_moments(b->d.precision, j.d.q, view, order3 ? prev : cur, R, R + n, R + 2 * n,
                                    R + 3 * n, b->stream);
    }
    // the density the last step took its force from: rho0 before any step
    // and, in the LBfluid2D order, for the first step; otherwise the moments
    // of f(N-1) (in the scmp3D order the r array itself)
    const bool force_rho0 = (done == 0) || (!order3 && done == 1);
    if (e == hipSuccess) {
        if (force_rho0) e = lbm::batch::widen(j.rho0, S, n, b->stream);
        else if (order3) e = hipMemcpyAsync(S, R, n * 8, hipMemcpyDeviceToDevice, b->stream);
        else e = lbm::batch::job_moments(b->d.precision, j.d.q, view, prev, S, S + n, S + 2 * n, S + 3 * n,
                                         b->stream);
    }
    if (e == hipSuccess) e = lbm::fp64_psi_from_rho(S, psi, n, b->stream);
    if (e == hipSuccess) {
        lbm::StepArgs a{};
        a.q = j.d.q;
        a.walls = 0;
        a.tau = j.d.tau;
        a.G = j.d.G;
        lbm::Grid g{j.d.nx, j.d.ny, j.d.q == 9 ? 1 : j.d.nz, n, n};
        e = lbm::fp64_force(a, g, psi, F, F + n, F + 2 * n, b->stream);
    }
    double* outs[7] = {rho, u, v, w, fx, fy, fz};
    const double* srcs[7] = {R, R + n, R + 2 * n, R + 3 * n, F, F + n, F + 2 * n};
    for (int k = 0; k < 7 && e == hipSuccess; k++)
        if (outs[k]) e = hipMemcpyAsync(outs[k], srcs[k], n * 8, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return lbm::set_error(LBM_EHIP, "read: %s", hipGetErrorString(e));
    return LBM_OK;
}

int lbm_batch_extend(lbm_batch* b, int64_t job_id, int64_t nsteps)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_extend: NULL batch");
    if (b->tick_pending) return in_flight("lbm_batch_extend");
    if (nsteps < 1) return lbm::set_error(LBM_EINVAL, "lbm_batch_extend: nsteps %lld < 1", (long long)nsteps);
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) return lbm::set_error(LBM_EINVAL, "unknown job %lld", (long long)job_id);
    HostJob* j = it->second.get();
    if (j->failed)
        return lbm::set_error(LBM_EJOB, "job %lld failed after step %lld and cannot continue", (long long)job_id,
                              (long long)j->done);
    if (!j->finished)
        return lbm::set_error(LBM_ESTATE, "job %lld is still running (%lld steps left)", (long long)job_id,
                              (long long)j->remaining);
    j->remaining = nsteps;
    j->d.steps = j->done + nsteps;
    j->finished = false;
    j->waited = 0;
    b->order.clear();
    b->dirty = true;   // the job rejoins the work tables at the next tick
    return LBM_OK;
}

int lbm_batch_job_status(const lbm_batch* b, int64_t job_id, lbm_job_status* st)
{
    if (!b || !st) return lbm::set_error(LBM_EINVAL, "lbm_batch_job_status: NULL argument");
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) return lbm::set_error(LBM_EINVAL, "unknown job %lld", (long long)job_id);
    const HostJob& j = *it->second;
    st->state = j.failed ? LBM_JOB_FAILED : (j.finished ? LBM_JOB_DONE : LBM_JOB_RUNNING);
    st->steps_done = j.done;
    st->steps_finite = j.steps_finite;
    st->mass = j.mass_valid ? j.mass : std::nan("");
    return LBM_OK;
}

int lbm_batch_cancel(lbm_batch* b, int64_t job_id)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_cancel: NULL batch");
    if (b->tick_pending) return in_flight("lbm_batch_cancel");
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) return lbm::set_error(LBM_EINVAL, "unknown job %lld", (long long)job_id);
    // ticks are synchronous, so no launch of this batch is in flight here
    free_job(b, it->second.get());
    b->jobs.erase(it);
    b->order.clear();
    b->dirty = true;   // the job leaves the work tables before the next tick
    return LBM_OK;
}

int lbm_batch_active(const lbm_batch* b) { return b ? static_cast<int>(b->jobs.size()) : -1; }

int64_t lbm_batch_bytes_in_use(const lbm_batch* b) { return b ? b->in_use : -1; }

int64_t lbm_batch_steps_run(const lbm_batch* b) { return b ? b->steps_run : -1; }

int lbm_batch_set_work_limit(lbm_batch* b, int64_t max_nodes)
{
    if (!b || max_nodes < 0) return lbm::set_error(LBM_EINVAL, "lbm_batch_set_work_limit: bad argument");
    if (b->tick_pending) return in_flight("lbm_batch_set_work_limit");
    b->work_limit = max_nodes;
    b->dirty = true;
    return LBM_OK;
}

int lbm_batch_set_max_wait(lbm_batch* b, int ticks)
{
    if (!b || ticks < 0) return lbm::set_error(LBM_EINVAL, "lbm_batch_set_max_wait: bad argument");
    if (b->tick_pending) return in_flight("lbm_batch_set_max_wait");
    b->max_wait = ticks;
    b->dirty = true;
    return LBM_OK;
}

int lbm_batch_set_concurrency(lbm_batch* b, int on)
{
    if (!b || (on != 0 && on != 1)) return lbm::set_error(LBM_EINVAL, "lbm_batch_set_concurrency: bad argument");
    if (b->tick_pending) return in_flight("lbm_batch_set_concurrency");
    if (b->concurrent != (on == 1)) {
        b->graph_valid = false;   // the captured launch sequence changes
        b->dirty = true;          // and the row items (row_chunks)
    }
    b->concurrent = on == 1;
    return LBM_OK;
}

int lbm_batch_set_launch(lbm_batch* b, int d2q9_chunk, int d2q9_rows_kernel, int d3q19_march)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_set_launch: NULL batch");
    if (b->tick_pending) return in_flight("lbm_batch_set_launch");
    if (d2q9_chunk < 1) return lbm::set_error(LBM_EINVAL, "lbm_batch_set_launch: chunk must be >= 1");
    if (d2q9_rows_kernel < 0 || d2q9_rows_kernel > 2)
        return lbm::set_error(LBM_EINVAL, "lbm_batch_set_launch: d2q9_rows_kernel must be 0, 1 or 2");
    b->chunk = d2q9_chunk;
    b->rows_kernel = d2q9_rows_kernel;
    if (d3q19_march < 0 || d3q19_march > 2)
        return lbm::set_error(LBM_EINVAL, "lbm_batch_set_launch: d3q19_march must be 0, 1 or 2");
    b->march19 = d3q19_march;
    b->tune_left = 0;
    b->dirty = true;
    b->graph_valid = false;
    return LBM_OK;
}

int lbm_batch_get_launch(const lbm_batch* b, int* d2q9_chunk, int* d2q9_rows_kernel, int* d3q19_march)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_get_launch: NULL batch");
    if (d2q9_chunk) *d2q9_chunk = b->chunk;
    if (d2q9_rows_kernel) *d2q9_rows_kernel = b->rows_kernel;
    if (d3q19_march) *d3q19_march = b->march19;
    return LBM_OK;
}

int lbm_batch_tuning(const lbm_batch* b, int* d2q9_chunk, int* d2q9_rows_kernel)
{
    if (!b) return lbm::set_error(LBM_EINVAL, "lbm_batch_tuning: NULL batch");
    if (d2q9_chunk) *d2q9_chunk = b->chunk;
    if (d2q9_rows_kernel) *d2q9_rows_kernel = b->rows_kernel;
    return LBM_OK;
}

}  // extern "C"
