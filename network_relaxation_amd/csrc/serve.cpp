// Serving loop of libnrelax.so: continuous batching of relax jobs on one GPU.
//
// Jobs (a network, its stiffnesses, FIRE parameters and a list of strain
// increments) are laid out on host worker threads, admitted into a device
// arena sized by an HBM budget as memory and slots free up, and relaxed by
// repeated launches of the persistent relax kernel.  Every launch gives each
// resident job at most `slice_steps` FIRE steps; a job's state (positions,
// velocities, forces and the FIRE scalars) stays on the device between
// launches and between its strain increments, which the kernel walks by
// itself.  Between two launches the host retires finished jobs and admits
// waiting ones, so work that arrives while others run overlaps with them on
// the device.  A job's arithmetic does not depend on what else is resident
// (team size is a function of the job and the server's shape), so its
// result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/nrelax.h"
#include "device.hpp"
#include "internal.hpp"
#include "plan.hpp"
#include "teams.hpp"

namespace {

using nr::DevBuf;
using nr::fail;
using nr::hip_fail;

// First-fit allocator of ranges of [0, cap) with coalescing frees.
class Ranges {
public:
    void reset(int64_t cap)
    {
        free_.clear();
        cap_ = cap;
        if (cap > 0) free_[0] = cap;
    }
    int64_t alloc(int64_t n)
    {
        if (n <= 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const int64_t off = it->first, size = it->second;
            free_.erase(it);
            if (size > n) free_[off + n] = size - n;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t n)
    {
        if (n <= 0) return;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t free_total() const
    {
        int64_t t = 0;
        for (auto &kv : free_) t += kv.second;
        return t;
    }
    int64_t cap() const { return cap_; }

private:
    std::map<int64_t, int64_t> free_;
    int64_t cap_ = 0;
};

enum JobState { kWaiting, kActive, kDone };

struct Job {
    int id = 0;
    nr::Network net;
    double ks = 0.0, kb = 0.0;
    nr_relax_params prm{};
    std::vector<double2> incs;
    bool keep_pos = false;
    double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
    // layout, built by a planner thread: 0 pending, 1 ready, 2 failed
    std::atomic<int> planned{0};
    nr::Plan plan;
    int plan_rc = 0;
    std::string plan_err;
    // placement
    JobState state = kWaiting;
    int slot = -1, team = 1;
    int64_t node_off = -1, tile_off = -1, blob_off = -1, inc_off = -1;
    int64_t blob16 = 0;
    // results
    int status = NR_CONVERGED, error = NR_OK, steps = 0, n_done = 0;
    std::vector<nr_report> reports;
    std::vector<double> pos;
    double an[4] = {0.0, 0.0, 0.0, 0.0};  // post-relax reductions (nr_job_result)
};

// Device tables of one launch of a slice (reused across slices).
struct Tables {
    DevBuf<nr::BlockRole> roles;
    DevBuf<int> class_jobs, class_off, queues, mailbox;
    DevBuf<unsigned long long> partials;
    DevBuf<unsigned long long> bar;
    std::vector<nr::BlockRole> h_roles;  // host copies, alive until the launch ends
    std::vector<int> h_jobs, h_off, h_slice;
    int grid = 0, n_teams = 0;  // n_teams: barrier/mailbox ids in use
    bool cooperative = false;
};

constexpr int kMaxClasses = 64;

}  // namespace

struct nr_server {
    int device = 0, precision = NR_PREC_FP64, block = 256, npt = 1, tile_nodes = 256;
    int team_cap = 256, tiles_per_member = 1, slice_steps = 2000, stall_steps = 2000;
    double switch_ratio = 3e-2;
    int64_t lds_budget = 80 * 1024;
    size_t lds = 0;
    int resident = 0;
    // arena
    int64_t node_cap = 0, tile_cap = 0, blob_cap = 0, inc_cap = 0;
    int slot_cap = 0;
    DevBuf<double2> x[2], v[2], f[2];
    DevBuf<float2> u32[2], v32[2], f32[2];
    DevBuf<unsigned char> mask;
    DevBuf<uint4> blob;
    DevBuf<nr::TileDev> tiles;
    DevBuf<nr::NetDev> dnet;
    DevBuf<nr::NetOut> dout;
    DevBuf<nr::NetRun> run;
    DevBuf<double2> incs;
    DevBuf<nr::NetOut> inc_out;
    DevBuf<unsigned> err;
    Ranges r_nodes, r_tiles, r_blob, r_incs;
    std::vector<int> free_slots;
    std::vector<int> slot_job;  // job id per slot, -1 = free
    std::vector<std::unique_ptr<Tables>> tables;
    hipStream_t stream = nullptr, ustream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool inflight = false;
    std::vector<int> inflight_slots;
    // jobs
    int next_id = 0;
    std::map<int, std::unique_ptr<Job>> jobs;
    std::deque<int> waiting;  // arrival order
    std::deque<int> done;     // finished, result not yet taken
    // planner pool
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job *> todo;
    bool stop = false;
    // statistics
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    int64_t slices = 0;
    double busy_ms = 0.0;

    double clock() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void planner()
    {
        for (;;) {
            Job *j = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !todo.empty(); });
                if (stop) return;
                j = todo.front();
                todo.pop_front();
            }
            std::vector<const nr::Network *> nets{&j->net};
            std::vector<double> ks{j->ks}, kb{j->kb};
            j->plan_rc = nr::build_plan(nets, ks, kb, tile_nodes, lds_budget, j->plan, j->plan_err);
            if (!j->plan_rc && j->plan.lds_bytes() > lds_budget) {
                j->plan_rc = NR_E_TOPO;
                j->plan_err = "tile working set exceeds the server's LDS size";
            }
            j->planned.store(j->plan_rc ? 2 : 1, std::memory_order_release);
        }
    }

    ~nr_server()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto &t : workers) t.join();
        if (inflight) (void)hipStreamSynchronize(stream);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
        if (ustream) (void)hipStreamDestroy(ustream);
    }
};

namespace {

double expected_steps(const Job &j, int slice_steps)
{
    // expected steps of the job's increments; a launch gives at most
    // slice_steps
    const double s = nr::expected_steps(j.net.n, j.plan.nets[0].n_bonds) * (double)j.incs.size();
    return std::min(s, (double)slice_steps);
}

// Cost of one FIRE step of a job in tile passes (the busiest member's tile
// passes and the team's exchange; teams.hpp's job_cost of one step).
double step_cost(const Job &j)
{
    return nr::job_cost(1.0, (int)j.plan.tiles.size(), std::max(1, j.team));
}

// Steps of each active job in the next launch: the slowest job per step gets
// slice_steps, the others as many more as fit in the same time (at most
// kSliceSpread times as many), so that a small job in a mix with a large one
// advances at the pace it would alone instead of idling its workgroups until
// the large job's slice ends.
constexpr int kSliceSpread = 64;

void release_job(nr_server *s, Job &j)
{
    s->r_nodes.release(j.node_off, j.net.n);
    s->r_tiles.release(j.tile_off, (int64_t)j.plan.tiles.size());
    s->r_blob.release(j.blob_off, j.blob16);
    s->r_incs.release(j.inc_off, (int64_t)j.incs.size());
    s->slot_job[j.slot] = -1;
    s->free_slots.push_back(j.slot);
    j.slot = -1;
    j.plan = nr::Plan();  // the device copy was the only use
}

// Places a laid-out job in the arena and uploads it on the upload stream.
// Returns 1 when it does not fit now, 0 when placed, an NR_E_ code on error.
int place_job(nr_server *s, Job &j)
{
    const nr::Plan &P = j.plan;
    const int n = j.net.n, nt = (int)P.tiles.size(), ninc = (int)j.incs.size();
    const int64_t b16 = (int64_t)(P.blob.size() / 16);
    if (s->free_slots.empty()) return 1;
    const int64_t no = s->r_nodes.alloc(n);
    if (no < 0) return 1;
    const int64_t to = s->r_tiles.alloc(nt);
    if (to < 0) {
        s->r_nodes.release(no, n);
        return 1;
    }
    const int64_t bo = s->r_blob.alloc(b16);
    if (bo < 0) {
        s->r_nodes.release(no, n);
        s->r_tiles.release(to, nt);
        return 1;
    }
    const int64_t io = s->r_incs.alloc(ninc);
    if (io < 0) {
        s->r_nodes.release(no, n);
        s->r_tiles.release(to, nt);
        s->r_blob.release(bo, b16);
        return 1;
    }
    const int slot = s->free_slots.back();
    s->free_slots.pop_back();
    s->slot_job[slot] = j.id;
    j.slot = slot;
    j.node_off = no;
    j.tile_off = to;
    j.blob_off = bo;
    j.inc_off = io;
    j.blob16 = b16;
    j.team = nr::team_size(nt, s->team_cap, s->tiles_per_member);

    // relocate the one-network layout to its arena offsets
    std::vector<nr::TileDev> td(P.tiles);
    std::vector<uint8_t> blob(P.blob);
    for (nr::TileDev &t : td) {
        const nr::BlobLayout Lb = nr::blob_layout(t.n_own, t.n_halo, t.nb, t.nt, t.n_inc);
        int32_t *halo = reinterpret_cast<int32_t *>(blob.data() + (size_t)t.blob16 * 16 + Lb.halo);
        for (int q = 0; q < t.n_halo; q++) halo[q] += (int32_t)no;
        t.net = slot;
        t.node0 += (int)no;
        t.blob16 += (int)bo;
    }
    const nr::PlanNet &pn = P.nets[0];
    std::vector<double2> x0(n);
    for (int d = 0; d < n; d++) x0[d] = make_double2(j.net.pos[2 * pn.perm[d]], j.net.pos[2 * pn.perm[d] + 1]);
    nr::NetDev nd;
    std::memset(&nd, 0, sizeof(nd));
    nd.node_base = (int)no;
    nd.n_nodes = n;
    nd.tile_base = (int)to;
    nd.n_tiles = nt;
    nd.width = j.net.width;
    nd.height = j.net.height;
    nd.miny = j.net.miny;
    nd.yrange = j.net.height > 0 ? j.net.height : j.net.maxy - j.net.miny;
    nd.ks = j.ks;
    nd.kb = j.kb;
    nd.a0 = j.prm.alpha0;
    nd.fa = j.prm.f_alpha;
    nd.dtmax = j.prm.dt_max;
    nd.finc = j.prm.f_inc;
    nd.fdec = j.prm.f_dec;
    nd.nmin = j.prm.nmin;
    nd.mass = j.prm.mass;
    nd.fcut = j.prm.fcut;
    nd.max_steps = j.prm.max_steps;
    nd.crit = j.prm.criterion;
    nd.method = j.prm.method;
    nd.inc_count = ninc;
    nd.inc_base = (int)io;
    nd.out_base = (int)io;

    hipStream_t u = s->ustream;
    HIP_TRY(hipMemcpyAsync(s->tiles.p + to, td.data(), sizeof(nr::TileDev) * nt,
                           hipMemcpyHostToDevice, u));
    HIP_TRY(hipMemcpyAsync(s->blob.p + bo, blob.data(), blob.size(), hipMemcpyHostToDevice, u));
    HIP_TRY(hipMemcpyAsync(s->mask.p + no, P.mask.data(), n, hipMemcpyHostToDevice, u));
    HIP_TRY(hipMemcpyAsync(s->x[0].p + no, x0.data(), sizeof(double2) * n,
                           hipMemcpyHostToDevice, u));
    for (int q = 0; q < 2; q++) {
        HIP_TRY(hipMemsetAsync(s->v[q].p + no, 0, sizeof(double2) * n, u));
        HIP_TRY(hipMemsetAsync(s->f[q].p + no, 0, sizeof(double2) * n, u));
        if (s->precision == NR_PREC_MIXED) {
            HIP_TRY(hipMemsetAsync(s->u32[q].p + no, 0, sizeof(float2) * n, u));
            HIP_TRY(hipMemsetAsync(s->v32[q].p + no, 0, sizeof(float2) * n, u));
            HIP_TRY(hipMemsetAsync(s->f32[q].p + no, 0, sizeof(float2) * n, u));
        }
    }
    HIP_TRY(hipMemcpyAsync(s->incs.p + io, j.incs.data(), sizeof(double2) * ninc,
                           hipMemcpyHostToDevice, u));
    HIP_TRY(hipMemcpyAsync(s->dnet.p + slot, &nd, sizeof(nd), hipMemcpyHostToDevice, u));
    HIP_TRY(hipMemsetAsync(s->run.p + slot, 0, sizeof(nr::NetRun), u));
    // the host arrays above are locals: finish the copies before they go
    HIP_TRY(hipStreamSynchronize(u));
    // only the tile count and node order are needed from here on
    j.plan.blob.clear();
    j.plan.blob.shrink_to_fit();
    j.plan.mask.clear();
    j.plan.mask.shrink_to_fit();
    return 0;
}

// True when the job can never fit in the arena, even empty.
bool never_fits(const nr_server *s, const Job &j)
{
    return j.net.n > s->r_nodes.cap() || (int64_t)j.plan.tiles.size() > s->r_tiles.cap() ||
           (int64_t)(j.plan.blob.size() / 16) > s->r_blob.cap() ||
           (int64_t)j.incs.size() > s->r_incs.cap();
}

// Admits waiting jobs in arrival order; a job that does not fit yet lets
// later (smaller) ones in behind it unless it has waited over a second.
int admit(nr_server *s)
{
    const double now = s->clock();
    bool blocked = false;
    for (auto it = s->waiting.begin(); it != s->waiting.end();) {
        Job &j = *s->jobs[*it];
        const int pl = j.planned.load(std::memory_order_acquire);
        if (pl == 0) {
            ++it;
            continue;
        }
        if (pl == 2 || never_fits(s, j)) {
            if (pl != 2) {
                j.plan_rc = NR_E_ALLOC;
                j.plan = nr::Plan();
            }
            j.state = kDone;
            j.status = NR_ABORTED;
            j.error = j.plan_rc;
            j.t_done = now;
            s->done.push_back(j.id);
            it = s->waiting.erase(it);
            continue;
        }
        if (blocked) break;
        const int rc = place_job(s, j);
        if (rc < 0 || rc > 1) return rc;
        if (rc == 1) {
            if (now - j.t_submit > 1.0) blocked = true;
            ++it;
            continue;
        }
        j.state = kActive;
        j.t_admit = now;
        it = s->waiting.erase(it);
    }
    return NR_OK;
}

int launch_slice(nr_server *s)
{
    std::vector<int> act;
    for (int k = 0; k < s->slot_cap; k++)
        if (s->slot_job[k] >= 0 && s->jobs[s->slot_job[k]]->state == kActive) act.push_back(k);
    s->inflight_slots = act;
    if (act.empty()) return NR_OK;
    std::vector<nr::JobSpec> specs(act.size());
    std::vector<int> budget(act.size(), s->slice_steps);
    double c_max = 0.0;
    for (int k : act) c_max = std::max(c_max, step_cost(*s->jobs[s->slot_job[k]]));
    for (size_t i = 0; i < act.size(); i++) {
        const Job &j = *s->jobs[s->slot_job[act[i]]];
        const double r = c_max / std::max(step_cost(j), 1e-9);
        budget[i] = (int)std::min<double>((double)s->slice_steps * kSliceSpread,
                                          std::max<double>(s->slice_steps,
                                                           std::floor(s->slice_steps * r)));
        specs[i] = {j.team, (int)j.plan.tiles.size(), expected_steps(j, budget[i])};
    }
    std::vector<nr::TeamLayout> layouts = nr::plan_launches(specs, s->resident);
    while (s->tables.size() < layouts.size()) {
        auto t = std::make_unique<Tables>();
        const int R = s->resident;
        HIP_TRY(t->roles.alloc(R));
        HIP_TRY(t->class_jobs.alloc(s->slot_cap));
        HIP_TRY(t->class_off.alloc(kMaxClasses + 1));
        HIP_TRY(t->queues.alloc(32 * kMaxClasses));
        // split trees: at most 2g - 1 ids per team of g, so 2R ids in all
        HIP_TRY(t->mailbox.alloc((size_t)32 * 2 * R));
        HIP_TRY(t->partials.alloc((size_t)nr::kPartWords * R));
        HIP_TRY(t->bar.alloc((size_t)nr::kBarStride * 2 * R));
        s->tables.push_back(std::move(t));
    }
    nr::KArgs a;
    std::memset(&a, 0, sizeof(a));
    for (int q = 0; q < 2; q++) {
        a.x[q] = s->x[q].p;
        a.v[q] = s->v[q].p;
        a.f[q] = s->f[q].p;
        a.u32[q] = s->u32[q].p;
        a.v32[q] = s->v32[q].p;
        a.f32[q] = s->f32[q].p;
    }
    a.mixed = s->precision == NR_PREC_MIXED ? 1 : 0;
    a.switch_ratio = s->switch_ratio;
    a.stall_steps = s->stall_steps;
    a.mask = s->mask.p;
    a.blob = s->blob.p;
    a.tiles = s->tiles.p;
    a.nets = s->dnet.p;
    a.out = s->dout.p;
    a.tile_cap = (int)(s->lds - nr::kCacheBytes);
    a.mode = nr::kModeRelax;
    a.err = s->err.p;
    a.timeout_ticks = 100000000ll * 60;
    a.run = s->run.p;
    a.incs = s->incs.p;
    a.inc_out = s->inc_out.p;
    a.slice_steps = s->slice_steps;
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    for (size_t l = 0; l < layouts.size(); l++) {
        nr::TeamLayout &L = layouts[l];
        Tables &t = *s->tables[l];
        const int ncls = (int)L.class_off.size() - 1;
        if (ncls > kMaxClasses) return fail(NR_E_ARG, "too many team-size classes");
        t.h_roles = L.roles;
        t.h_jobs.resize(L.class_jobs.size());
        for (size_t i = 0; i < L.class_jobs.size(); i++) t.h_jobs[i] = act[L.class_jobs[i]];
        t.h_off = L.class_off;
        t.h_slice.assign(act.size(), 0);
        for (size_t i = 0; i < act.size(); i++) t.h_slice[i] = budget[i];
        t.grid = L.grid;
        t.n_teams = L.n_ids;
        t.cooperative = L.cooperative;
        HIP_TRY(hipMemcpyAsync(t.roles.p, t.h_roles.data(), sizeof(nr::BlockRole) * t.h_roles.size(),
                               hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(t.class_jobs.p, t.h_jobs.data(), sizeof(int) * t.h_jobs.size(),
                               hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(t.class_off.p, t.h_off.data(), sizeof(int) * t.h_off.size(),
                               hipMemcpyHostToDevice, s->stream));
        // each active job's steps of this launch, into its resumable state
        for (size_t i = 0; i < act.size(); i++)
            HIP_TRY(hipMemcpyAsync(&(s->run.p + act[i])->slice, &t.h_slice[i], sizeof(int),
                                   hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemsetAsync(t.queues.p, 0, sizeof(int) * 32 * ncls, s->stream));
        // partial-sum records carry per-launch step tags starting at 1
        HIP_TRY(hipMemsetAsync(t.partials.p, 0, sizeof(unsigned long long) * t.partials.n,
                               s->stream));
        // team compositions change between slices: restart the barrier counters
        HIP_TRY(hipMemsetAsync(t.bar.p, 0,
                               sizeof(unsigned long long) * nr::kBarStride * std::max(1, t.n_teams),
                               s->stream));
        a.roles = t.roles.p;
        a.class_jobs = t.class_jobs.p;
        a.class_off = t.class_off.p;
        a.queues = t.queues.p;
        a.partials = t.partials.p;
        a.bar = t.bar.p;
        a.mailbox = t.mailbox.p;
        std::memcpy(a.cls_of_log2g, L.cls_of_log2g, sizeof(a.cls_of_log2g));
        int e = nr::launch_relax(a, s->block, s->npt, NR_METHOD_FIRE, t.grid, s->lds,
                                 t.cooperative, s->stream);
        if (e) return hip_fail((hipError_t)e, "relax kernel launch");
    }
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    s->inflight = true;
    s->slices++;
    return NR_OK;
}

// Post-relax reductions of a finished job (as nr_batch_analyze) from its
// resident state: the input geometry and the bond list go up, the reductions
// come back.
int analyze_job(nr_server *s, Job &j, const nr::NetDev &nd)
{
    const nr::Network &N = j.net;
    const nr::PlanNet &pn = j.plan.nets[0];
    const int ne = N.e();
    std::vector<int4> bonds(std::max(ne, 1));
    for (int e = 0; e < ne; e++)
        bonds[e] = make_int4(pn.iperm[N.bonds[2 * e]], pn.iperm[N.bonds[2 * e + 1]],
                             N.bwrap[2 * e], N.bwrap[2 * e + 1]);
    std::vector<double2> x0(std::max(N.n, 1));
    for (int d = 0; d < N.n; d++)
        x0[d] = make_double2(N.pos[2 * pn.perm[d]], N.pos[2 * pn.perm[d] + 1]);
    DevBuf<int4> d_bonds;
    DevBuf<double2> d_x0;
    DevBuf<double> strain, part;
    HIP_TRY(d_bonds.upload(bonds));
    HIP_TRY(d_x0.upload(x0));
    HIP_TRY(strain.alloc(std::max(ne, 1)));
    HIP_TRY(part.alloc(4 * nr::kAnalysisBlocks));
    const double yr = N.height > 0 ? N.height : N.maxy - N.miny;
    double gx = 0.0, gy = 0.0;
    for (int i = 0; i < j.n_done; i++) {
        gx += j.incs[i].x;
        gy += j.incs[i].y;
    }
    HIP_TRY((hipError_t)nr::launch_analysis(s->x[nd.parity].p + j.node_off, d_x0.p, N.n, N.miny,
                                            gx / yr, gy / yr, d_bonds.p, ne, N.width, N.height,
                                            nd.le_shift, strain.p, part.p, s->stream));
    std::vector<double> hp(4 * nr::kAnalysisBlocks);
    HIP_TRY(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    double na = 0.0, ax = 0.0, sum = 0.0, mx = 0.0;
    for (int w = 0; w < nr::kAnalysisBlocks; w++) {
        na += hp[w];
        ax += hp[nr::kAnalysisBlocks + 3 * w];
        sum += hp[nr::kAnalysisBlocks + 3 * w + 1];
        mx = std::max(mx, hp[nr::kAnalysisBlocks + 3 * w + 2]);
    }
    j.an[0] = N.n > 0 ? na / N.n : 0.0;
    j.an[1] = j.ks * ax;
    j.an[2] = ne > 0 ? sum / ne : 0.0;
    j.an[3] = mx;
    return NR_OK;
}

// Waits for the launch in flight and retires the jobs it finished.
int retire(nr_server *s)
{
    if (!s->inflight) return NR_OK;
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->inflight = false;
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->busy_ms += ms;
    unsigned err = 0;
    HIP_TRY(hipMemcpy(&err, s->err.p, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (err) {
        (void)hipMemset(s->err.p, 0, sizeof(unsigned));
        return fail(NR_E_TIMEOUT, "a team barrier exceeded its wait bound");
    }
    if (s->inflight_slots.empty()) return NR_OK;
    std::vector<nr::NetOut> out(s->slot_cap);
    HIP_TRY(hipMemcpy(out.data(), s->dout.p, sizeof(nr::NetOut) * s->slot_cap,
                      hipMemcpyDeviceToHost));
    const double now = s->clock();
    for (int slot : s->inflight_slots) {
        const nr::NetOut &o = out[slot];
        if (o.status == NR_IN_PROGRESS) continue;
        Job &j = *s->jobs[s->slot_job[slot]];
        nr::NetRun run;
        nr::NetDev nd;
        HIP_TRY(hipMemcpy(&run, s->run.p + slot, sizeof(run), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&nd, s->dnet.p + slot, sizeof(nd), hipMemcpyDeviceToHost));
        const int ninc = std::min<int>(run.inc_next, (int)j.incs.size());
        std::vector<nr::NetOut> io(ninc);
        if (ninc > 0)
            HIP_TRY(hipMemcpy(io.data(), s->inc_out.p + j.inc_off, sizeof(nr::NetOut) * ninc,
                              hipMemcpyDeviceToHost));
        j.reports.resize(ninc);
        j.status = NR_CONVERGED;
        j.steps = 0;
        for (int i = 0; i < ninc; i++) {
            nr::to_report(io[i], j.reports[i]);
            j.steps += io[i].steps;
            if (j.status == NR_CONVERGED && io[i].status != NR_CONVERGED) j.status = io[i].status;
        }
        if (ninc == 0) j.status = o.status;
        j.n_done = ninc;
        if (j.keep_pos) {
            std::vector<double2> xs(j.net.n);
            HIP_TRY(hipMemcpy(xs.data(), s->x[nd.parity].p + j.node_off,
                              sizeof(double2) * j.net.n, hipMemcpyDeviceToHost));
            j.pos.assign(2 * (size_t)j.net.n, 0.0);
            const std::vector<int> &perm = j.plan.nets[0].perm;
            for (int d = 0; d < j.net.n; d++) {
                j.pos[2 * perm[d]] = xs[d].x;
                j.pos[2 * perm[d] + 1] = xs[d].y;
            }
        }
        const int arc = analyze_job(s, j, nd);
        if (arc) return arc;
        j.t_done = now;
        j.state = kDone;
        release_job(s, j);
        s->done.push_back(j.id);
    }
    return NR_OK;
}

}  // namespace

extern "C" {

int nr_server_create(const nr_batch_opts *shape, const nr_server_opts *opts, nr_server **out)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!out) return fail(NR_E_ARG, "null argument");
        auto s = std::make_unique<nr_server>();
        s->device = shape ? shape->device : 0;
        s->precision = shape ? shape->precision : NR_PREC_FP64;
        if (s->precision != NR_PREC_FP64 && s->precision != NR_PREC_MIXED)
            return fail(NR_E_ARG, "unknown precision");
        if (shape && shape->mixed_switch > 0) s->switch_ratio = shape->mixed_switch;
        s->block = shape && shape->block > 0 ? shape->block : 256;
        if (s->block != 256 && s->block != 512) return fail(NR_E_ARG, "block must be 256 or 512");
        s->tile_nodes = shape && shape->tile_nodes > 0 ? shape->tile_nodes : s->block;
        if (s->tile_nodes > 2 * s->block) return fail(NR_E_ARG, "tile_nodes above 2 x block");
        if (s->precision == NR_PREC_MIXED && s->tile_nodes > s->block)
            return fail(NR_E_ARG, "mixed precision needs tile_nodes <= block");
        s->npt = s->tile_nodes > s->block ? 2 : 1;
        s->lds_budget = nr::lds_per_workgroup(s->block) - nr::kLdsReserve;
        s->lds = (size_t)(nr::kCacheBytes + s->lds_budget);
        // a served mix holds tens of jobs, not hundreds: the widest teams (one
        // tile per member) give the lowest latency and, measured, the highest
        // throughput there
        s->tiles_per_member = shape && shape->tiles_per_member > 0 ? shape->tiles_per_member : 1;
        const int64_t hbm = opts && opts->hbm_bytes > 0 ? opts->hbm_bytes : (int64_t)8 << 30;
        s->slot_cap = opts && opts->max_jobs > 0 ? opts->max_jobs : 4096;
        s->slice_steps = opts && opts->slice_steps > 0 ? opts->slice_steps : 2000;
        const int threads = opts && opts->planner_threads > 0 ? opts->planner_threads : 4;

        HIP_TRY(hipSetDevice(s->device));
        int per_cu = 0, n_cu = 0;
        HIP_TRY((hipError_t)nr::max_resident_blocks(s->block, s->npt, s->lds, &per_cu, &n_cu));
        if (per_cu < 1) return fail(NR_E_ARG, "kernel cannot be resident with this LDS size");
        s->resident = per_cu * n_cu;
        s->team_cap = std::min({shape && shape->team > 0 ? shape->team : 512, s->resident,
                                nr::max_team(s->block, s->npt)});

        // arena: per node the FP64 state (x, v, F in two parities), its mask, the
        // FP32 stage state when mixed, and about 88 B of tile topology
        const int64_t per_node = 97 + (s->precision == NR_PREC_MIXED ? 48 : 0) + 88;
        s->node_cap = std::min<int64_t>(hbm / per_node, 0x7fffffffll);
        s->blob_cap = std::min<int64_t>(s->node_cap * 88 / 16, 0x7fffffffll);
        s->tile_cap = s->node_cap / 64 + s->slot_cap;
        s->inc_cap = (int64_t)s->slot_cap * 32;
        if (s->node_cap < 256) return fail(NR_E_ARG, "HBM budget too small");
        const size_t N = (size_t)s->node_cap;
        for (int q = 0; q < 2; q++) {
            HIP_TRY(s->x[q].alloc(N));
            HIP_TRY(s->v[q].alloc(N));
            HIP_TRY(s->f[q].alloc(N));
            if (s->precision == NR_PREC_MIXED) {
                HIP_TRY(s->u32[q].alloc(N));
                HIP_TRY(s->v32[q].alloc(N));
                HIP_TRY(s->f32[q].alloc(N));
            }
        }
        HIP_TRY(s->mask.alloc(N));
        HIP_TRY(s->blob.alloc((size_t)s->blob_cap));
        HIP_TRY(s->tiles.alloc((size_t)s->tile_cap));
        HIP_TRY(s->dnet.alloc(s->slot_cap));
        HIP_TRY(s->dout.alloc(s->slot_cap));
        HIP_TRY(s->run.alloc(s->slot_cap));
        HIP_TRY(s->incs.alloc((size_t)s->inc_cap));
        HIP_TRY(s->inc_out.alloc((size_t)s->inc_cap));
        HIP_TRY(s->err.alloc(1));
        HIP_TRY(hipMemset(s->err.p, 0, sizeof(unsigned)));
        HIP_TRY(hipMemset(s->dout.p, 0, sizeof(nr::NetOut) * s->slot_cap));
        s->r_nodes.reset(s->node_cap);
        s->r_tiles.reset(s->tile_cap);
        s->r_blob.reset(s->blob_cap);
        s->r_incs.reset(s->inc_cap);
        s->slot_job.assign(s->slot_cap, -1);
        for (int k = s->slot_cap - 1; k >= 0; k--) s->free_slots.push_back(k);
        HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&s->ustream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&s->ev0));
        HIP_TRY(hipEventCreate(&s->ev1));
        nr_server *raw = s.get();
        for (int t = 0; t < threads; t++) s->workers.emplace_back([raw] { raw->planner(); });
        *out = s.release();
        return NR_OK;
    });
}

double nr_server_clock(const nr_server *s) { return s ? s->clock() : 0.0; }

int nr_server_submit(nr_server *s, const nr_network *net, double ks, double kb,
                     const nr_relax_params *params, int n_inc, const double *incs,
                     int keep_positions, int *job)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!s || !net || !params || n_inc < 1 || !incs) return fail(NR_E_ARG, "invalid arguments");
        if (!(ks >= 0) || !(kb >= 0)) return fail(NR_E_ARG, "negative stiffness");
        if (kb > 0 && !net->net.triples_built)
            return fail(NR_E_ARG, "bending stiffness > 0 but triples were not built");
        if (params->method != NR_METHOD_FIRE)
            return fail(NR_E_ARG, "the server relaxes by FIRE (CG runs through nr_batch_relax)");
        if (!(params->mass > 0) || params->max_steps < 0 || params->fcut < 0)
            return fail(NR_E_ARG, "invalid relax parameters");
        if (n_inc > 32 * 64) return fail(NR_E_ARG, "too many increments");
        for (int i = 0; i < n_inc; i++) {
            if (!std::isfinite(incs[2 * i]) || !std::isfinite(incs[2 * i + 1]))
                return fail(NR_E_ARG, "non-finite strain increment");
            if (net->net.height > 0 && incs[2 * i + 1] != 0.0)
                return fail(NR_E_ARG, "compression is not defined in the Lees-Edwards box");
        }
        auto j = std::make_unique<Job>();
        j->id = s->next_id++;
        j->net = net->net;
        j->ks = ks;
        j->kb = kb;
        j->prm = *params;
        j->incs.resize(n_inc);
        for (int i = 0; i < n_inc; i++) j->incs[i] = make_double2(incs[2 * i], incs[2 * i + 1]);
        j->keep_pos = keep_positions != 0;
        j->t_submit = s->clock();
        Job *raw = j.get();
        s->jobs[j->id] = std::move(j);
        s->waiting.push_back(raw->id);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->todo.push_back(raw);
        }
        s->cv.notify_one();
        if (job) *job = raw->id;
        return NR_OK;
    });
}

int nr_server_step(nr_server *s, int *n_active, int *n_waiting)
{
    if (!s) return fail(NR_E_ARG, "null server");
    HIP_TRY(hipSetDevice(s->device));
    int rc = retire(s);
    if (rc) return rc;
    rc = launch_slice(s);
    if (rc) return rc;
    // admissions overlap the launch just started; they join the next one
    rc = admit(s);
    if (rc) return rc;
    int act = 0;
    for (int k = 0; k < s->slot_cap; k++)
        if (s->slot_job[k] >= 0) act++;
    if (n_active) *n_active = act;
    if (n_waiting) *n_waiting = (int)s->waiting.size();
    return NR_OK;
}

int nr_server_drain(nr_server *s)
{
    if (!s) return fail(NR_E_ARG, "null server");
    for (;;) {
        int act = 0, wait = 0;
        int rc = nr_server_step(s, &act, &wait);
        if (rc) return rc;
        if (act == 0 && wait == 0 && !s->inflight) return NR_OK;
        if (act == 0 && !s->inflight) std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
}

int nr_server_next_done(nr_server *s, int *job)
{
    if (!s || !job) return fail(NR_E_ARG, "null argument");
    if (s->done.empty()) return 1;
    *job = s->done.front();
    return NR_OK;
}

int nr_server_result(nr_server *s, int job, nr_job_result *res, nr_report *reports, double *pos)
{
    if (!s) return fail(NR_E_ARG, "null server");
    auto it = s->jobs.find(job);
    if (it == s->jobs.end() || it->second->state != kDone)
        return fail(NR_E_ARG, "no finished job with this id");
    Job &j = *it->second;
    if (res) {
        std::memset(res, 0, sizeof(*res));
        res->job = j.id;
        res->status = j.status;
        res->error = j.error;
        res->n_inc = j.n_done;
        res->steps = j.steps;
        res->n_nodes = j.net.n;
        res->t_submit = j.t_submit;
        res->t_admit = j.t_admit;
        res->t_done = j.t_done;
        if (!j.reports.empty()) {
            res->e_stretch = j.reports.back().e_stretch;
            res->e_bend = j.reports.back().e_bend;
            res->max_force = j.reports.back().max_force;
        }
        res->nonaffinity = j.an[0];
        res->axial_energy = j.an[1];
        res->mean_bond_strain = j.an[2];
        res->max_bond_strain = j.an[3];
    }
    if (reports) std::copy(j.reports.begin(), j.reports.end(), reports);
    if (pos && !j.pos.empty()) std::copy(j.pos.begin(), j.pos.end(), pos);
    for (auto d = s->done.begin(); d != s->done.end(); ++d)
        if (*d == job) {
            s->done.erase(d);
            break;
        }
    s->jobs.erase(it);
    return NR_OK;
}

int nr_server_stats(const nr_server *s, nr_server_info *st)
{
    if (!s || !st) return fail(NR_E_ARG, "null argument");
    std::memset(st, 0, sizeof(*st));
    st->slices = s->slices;
    st->busy_ms = s->busy_ms;
    st->node_cap = s->node_cap;
    st->nodes_free = s->r_nodes.free_total();
    st->resident_workgroups = s->resident;
    for (int k = 0; k < s->slot_cap; k++)
        if (s->slot_job[k] >= 0) st->active++;
    st->waiting = (int)s->waiting.size();
    st->done = (int)s->done.size();
    return NR_OK;
}

void nr_server_destroy(nr_server *s) { delete s; }

}  // extern "C"
