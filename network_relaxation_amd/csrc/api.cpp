// C ABI of libnrelax.so: host networks, device batches and the launch of the
// persistent relax kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nrelax.h"
#include "device.hpp"
#include "internal.hpp"
#include "network.hpp"
#include "plan.hpp"
#include "teams.hpp"

namespace nr {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return NR_E_HIP;
}

void to_report(const NetOut &o, nr_report &r)
{
    r.rms_force = o.rms;
    r.max_force = o.fmax;
    r.e_stretch = o.es;
    r.e_bend = o.eb;
    r.steps = o.steps;
    r.status = o.status;
    r.fp64_steps = o.fp64_steps;
    r.pad = 0;
    r.t_start = (double)o.t_start * 1e-8;
    r.t_done = (double)o.t_done * 1e-8;
    r.shear_stress = o.sxy;
}

}  // namespace nr

namespace {

using nr::DevBuf;
using nr::fail;
using nr::hip_fail;

// Largest tiles per team member the default tuning considers: the best
// measured on MI355X for a batch of 512 networks of 1 to 150 tiles of 256
// nodes relaxed again (planned from the step counts of its last relax):
// 5 -> 79, 7 -> 81, 8 -> 84, 10 -> 80, 12 -> 65 networks/s (the cost model
// prefers 10, which measures slower).  A first relax, planned from estimated
// step counts, takes 6.5 s at 5 and 6.8 s at 8 (one badly estimated network
// ends last).
constexpr int kDefaultTilesPerMember = 8;
// Default largest team: every resident workgroup of an MI355X at two 256-thread
// workgroups per CU (a fixed number, so that team sizes do not depend on the
// batch; the resident count still bounds it).
constexpr int kDefaultTeamCap = 512;

// Device tables of one persistent launch.
struct LaunchDev {
    nr::TeamLayout L;
    DevBuf<nr::BlockRole> roles;
    DevBuf<int> class_jobs, class_off, queues, mailbox;
    DevBuf<unsigned long long> partials;
    DevBuf<unsigned long long> bar;
};

}  // namespace

struct nr_batch {
    int device = 0;
    int precision = NR_PREC_FP64;
    int npt = 1, block = 256, team_cap = 256, tiles_per_member = kDefaultTilesPerMember;
    int resident = 0;             // workgroups the relax kernel keeps resident
    std::vector<double> hist_steps;  // steps of each network in the last relax
    bool hist_planned = false;       // the current team layout used hist_steps
    int method = NR_METHOD_FIRE;  // of the last relax call (one method per call)
    size_t lds = 0;
    std::vector<std::unique_ptr<LaunchDev>> launches;
    std::vector<int> team_of_net;
    nr::Plan plan;
    // borrowed: read while the batch is created, and by nr_batch_analyze
    std::vector<const nr::Network *> nets;
    std::vector<nr::NetDev> hnet;
    std::vector<double> ks, kb;
    DevBuf<double2> x[2], v[2], f[2], x_init;
    DevBuf<float2> u32[2], v32[2], f32[2];  // mixed precision FP32 stage
    double switch_ratio = 3e-2;  // FP32 stage stops at this fraction of the initial |F|max
    int stall_steps = 2000;      // ... or after this many steps without a new minimum
    DevBuf<unsigned char> mask;
    DevBuf<uint4> blob;
    DevBuf<nr::TileDev> tiles;
    DevBuf<nr::NetDev> dnet;
    DevBuf<nr::NetOut> dout;
    DevBuf<nr::NetRun> run;        // resumable relax state, zeroed per call
    DevBuf<double2> incs;          // strain ramp increments
    DevBuf<nr::NetOut> inc_out;    // per-increment reports of a ramp
    DevBuf<unsigned> err;
    // post-relax reductions: per-network bond lists (device node indices),
    // built on first use
    std::vector<std::unique_ptr<DevBuf<int4>>> an_bonds;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    // captured strain step (nr_batch_graph_capture)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    DevBuf<double2> gstep;             // per-network affine step of the graph
    nr::NetOut *h_out = nullptr;       // pinned copies written by the graph
    nr::NetDev *h_net = nullptr;
    unsigned *h_err = nullptr;

    void drop_graph()
    {
        // a replay that returned early may still be queued on the stream
        if (gexec && stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    ~nr_batch()
    {
        drop_graph();
        if (h_out) (void)hipHostFree(h_out);
        if (h_net) (void)hipHostFree(h_net);
        if (h_err) (void)hipHostFree(h_err);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

extern "C" {

const char *nr_last_error(void) { return nr::g_err.c_str(); }

const char *nr_version(void) { return "nrelax 0.1 (gfx950)"; }

int nr_network_read(const char *path, nr_network **out)
{
    return nr::guarded(NR_E_OPEN, [&]() -> int {
        if (!path || !out) return fail(NR_E_ARG, "null argument");
        auto h = std::make_unique<nr_network>();
        std::string err;
        int rc = nr::read_network(path, h->net, err);
        if (rc) return fail(rc, err);
        *out = h.release();
        return NR_OK;
    });
}

int nr_network_from_arrays(int n_nodes, int n_bonds, double width, double height,
                           const double *pos, const int *bonds, const signed char *wraps,
                           nr_network **out)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!out) return fail(NR_E_ARG, "null argument");
        auto h = std::make_unique<nr_network>();
        std::string err;
        int rc = nr::network_from_arrays(n_nodes, n_bonds, width, height, pos, bonds, wraps,
                                         h->net, err);
        if (rc) return fail(rc, err);
        *out = h.release();
        return NR_OK;
    });
}

int nr_network_lattice(int L, double p, uint64_t seed, int lees_edwards, nr_network **out)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!out) return fail(NR_E_ARG, "null argument");
        auto h = std::make_unique<nr_network>();
        std::string err;
        int rc = nr::make_lattice(L, p, seed, lees_edwards != 0, h->net, err);
        if (rc) return fail(rc, err);
        *out = h.release();
        return NR_OK;
    });
}

int nr_network_build_triples(nr_network *net)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!net) return fail(NR_E_ARG, "null network");
        return nr::build_triples(net->net);
    });
}

int nr_network_set_fiber_kappa(nr_network *net, double lo, double hi, uint64_t seed)
{
    if (!net) return fail(NR_E_ARG, "null network");
    int rc = nr::set_fiber_kappa(net->net, lo, hi, seed);
    if (rc) return fail(rc, "fiber kappa needs built triples and 0 < lo <= hi");
    return NR_OK;
}

int nr_network_get_info(const nr_network *net, nr_network_info *info)
{
    if (!net || !info) return fail(NR_E_ARG, "null argument");
    info->n_nodes = net->net.n;
    info->n_bonds = net->net.e();
    info->n_triples = net->net.t();
    info->width = net->net.width;
    info->height = net->net.height;
    info->miny = net->net.miny;
    info->maxy = net->net.maxy;
    return NR_OK;
}

int nr_network_get_arrays(const nr_network *net, double *pos, int *bonds,
                          signed char *bond_wraps, double *rest, int *triples,
                          signed char *triple_wraps, double *triple_kappa,
                          unsigned char *fcodes)
{
    if (!net) return fail(NR_E_ARG, "null network");
    const nr::Network &n = net->net;
    if (pos) std::copy(n.pos.begin(), n.pos.end(), pos);
    if (bonds) std::copy(n.bonds.begin(), n.bonds.end(), bonds);
    if (bond_wraps) std::copy(n.bwrap.begin(), n.bwrap.end(), bond_wraps);
    if (rest) std::copy(n.rest.begin(), n.rest.end(), rest);
    if (triples) std::copy(n.tri.begin(), n.tri.end(), triples);
    if (triple_wraps) std::copy(n.twrap.begin(), n.twrap.end(), triple_wraps);
    if (triple_kappa) {
        if (n.tkb.empty()) std::fill(triple_kappa, triple_kappa + n.t(), 0.0);
        else std::copy(n.tkb.begin(), n.tkb.end(), triple_kappa);
    }
    if (fcodes) std::copy(n.fcode.begin(), n.fcode.end(), fcodes);
    return NR_OK;
}

int nr_network_write_dat(const nr_network *net, const char *path)
{
    if (!net || !path) return fail(NR_E_ARG, "null argument");
    int rc = nr::write_dat(net->net, path);
    if (rc) return fail(rc, std::string("cannot write ") + path);
    return NR_OK;
}

void nr_network_free(nr_network *net) { delete net; }

// ---- batches --------------------------------------------------------------

// Workgroup shape: `block` threads (256 or 512) each owning up to npt (1 or 2)
// nodes of a tile; default 256 threads, one node each (256-node tiles), two
// workgroups per CU (LDS budget 80 KB each).
static int choose_shapes(nr_batch *b, const nr_batch_opts *opts, int64_t *lds_budget)
{
    int block = (opts && opts->block > 0) ? opts->block : 256;
    if (block != 256 && block != 512) return fail(NR_E_ARG, "block must be 256 or 512");
    int tile = (opts && opts->tile_nodes > 0) ? opts->tile_nodes : block;
    if (tile > 2 * block) return fail(NR_E_ARG, "tile_nodes above 2 x block");
    if (b->precision == NR_PREC_MIXED && tile > block)
        return fail(NR_E_ARG, "mixed precision needs tile_nodes <= block");
    b->block = block;
    b->npt = tile > block ? 2 : 1;
    b->plan.tile_nodes = tile;
    // two workgroups per CU at 256 threads (one at 512): the tile area gets
    // what the LDS cache and the static scratch leave
    *lds_budget = nr::lds_per_workgroup(block) - nr::kLdsReserve;
    return NR_OK;
}

// Teams: each network gets a power-of-two team of workgroups (a function of
// its tile count and the options only, so its arithmetic does not depend on
// the rest of the batch); the resident workgroups are shared between the
// team-size classes to minimise the predicted finish time.
static int set_team_layout(nr_batch *b, int tiles_per_member)
{
    const nr::Plan &P = b->plan;
    const int count = (int)P.nets.size();
    b->tiles_per_member = tiles_per_member;
    std::vector<nr::JobSpec> jobs(count);
    for (int k = 0; k < count; k++)
        jobs[k] = {nr::team_size(P.nets[k].n_tiles, b->team_cap, b->tiles_per_member),
                   P.nets[k].n_tiles,
                   b->hist_steps.empty()
                       ? nr::expected_steps(P.nets[k].n_nodes, P.nets[k].n_bonds)
                       : b->hist_steps[k]};
    std::vector<nr::TeamLayout> layouts = nr::plan_launches(jobs, b->resident);
    b->launches.clear();
    b->team_of_net.assign(count, 1);
    for (nr::TeamLayout &L : layouts) {
        auto ld = std::make_unique<LaunchDev>();
        for (int k = 0; k < count; k++)
            if (L.team_of_net[k] > 0) b->team_of_net[k] = L.team_of_net[k];
        HIP_TRY(ld->roles.upload(L.roles));
        HIP_TRY(ld->class_jobs.upload(L.class_jobs));
        HIP_TRY(ld->class_off.upload(L.class_off));
        HIP_TRY(ld->queues.alloc((size_t)32 * (L.class_off.size() - 1)));
        HIP_TRY(ld->mailbox.alloc((size_t)32 * std::max(1, L.n_ids)));
        HIP_TRY(ld->partials.alloc((size_t)nr::kPartWords * std::max(1, L.n_parts)));
        HIP_TRY(ld->bar.alloc((size_t)nr::kBarStride * std::max(1, L.n_ids)));
        HIP_TRY(hipMemset(ld->bar.p, 0, sizeof(unsigned long long) * nr::kBarStride *
                                             std::max(1, L.n_ids)));
        ld->L = std::move(L);
        b->launches.push_back(std::move(ld));
    }
    b->drop_graph();  // a captured step refers to the old tables
    return NR_OK;
}

// Tiles per member when the options leave it open: the one of {1, 2, 4, 8}
// with the least predicted time for these networks (estimated step counts).
// A batch of many networks keeps every workgroup busy with small teams (8 is
// the best measured for 512 networks of 1 to 150 tiles); a few large networks
// get wide teams instead of leaving workgroups idle (a single network of 1024
// tiles: teams of 512 rather than 128).  The choice depends on the batch, so
// a network's reduction order may differ between batches; an explicit
// tiles_per_member fixes it.
static int tuned_tiles_per_member(const nr_batch *b)
{
    const nr::Plan &P = b->plan;
    std::vector<int> tiles(P.nets.size());
    std::vector<double> steps(P.nets.size());
    for (size_t k = 0; k < P.nets.size(); k++) {
        tiles[k] = P.nets[k].n_tiles;
        steps[k] = nr::expected_steps(P.nets[k].n_nodes, P.nets[k].n_bonds);
    }
    return nr::tune_tiles_per_member(tiles, steps, b->team_cap, b->resident,
                                     {1, 2, 4, kDefaultTilesPerMember}, nullptr);
}

int nr_batch_create(int count, nr_network *const *nets, const double *ks, const double *kb,
                    const nr_batch_opts *opts, nr_batch **out)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (count <= 0 || !nets || !ks || !kb || !out) return fail(NR_E_ARG, "invalid arguments");
        auto b = std::make_unique<nr_batch>();
        b->device = opts ? opts->device : 0;
        b->precision = opts ? opts->precision : NR_PREC_FP64;
        if (opts && opts->mixed_switch > 0) b->switch_ratio = opts->mixed_switch;
        if (b->precision != NR_PREC_FP64 && b->precision != NR_PREC_MIXED)
            return fail(NR_E_ARG, "unknown precision");
        HIP_TRY(hipSetDevice(b->device));
        int64_t lds_budget = 0;
        int rc = choose_shapes(b.get(), opts, &lds_budget);
        if (rc) return rc;
        for (int i = 0; i < count; i++) {
            if (!nets[i]) return fail(NR_E_ARG, "null network in batch");
            if (!(ks[i] >= 0) || !(kb[i] >= 0)) return fail(NR_E_ARG, "negative stiffness");
            b->nets.push_back(&nets[i]->net);
            b->ks.push_back(ks[i]);
            b->kb.push_back(kb[i]);
        }
        std::string err;
        rc = nr::build_plan(b->nets, b->ks, b->kb, b->plan.tile_nodes, lds_budget, b->plan, err);
        if (rc) return fail(rc, err);
        const nr::Plan &P = b->plan;
        b->lds = (size_t)(nr::kCacheBytes + P.lds_bytes());
        if (b->lds > 160 * 1024) return fail(NR_E_TOPO, "tile working set exceeds LDS");

        int per_cu = 0, n_cu = 0;
        HIP_TRY((hipError_t)nr::max_resident_blocks(b->block, b->npt, b->lds, &per_cu, &n_cu));
        if (per_cu < 1) return fail(NR_E_ARG, "kernel cannot be resident with this LDS size");
        // the whole budget, when that costs no residency: what the largest
        // tile leaves holds tile blobs resident across steps
        const size_t full = (size_t)(nr::kCacheBytes + lds_budget);
        if (full > b->lds) {
            int pc = 0, nc = 0;
            HIP_TRY((hipError_t)nr::max_resident_blocks(b->block, b->npt, full, &pc, &nc));
            if (pc == per_cu) b->lds = full;
        }
        b->resident = per_cu * n_cu;
        b->team_cap = std::min({opts && opts->team > 0 ? opts->team : kDefaultTeamCap, b->resident,
                                nr::max_team(b->block, b->npt)});
        rc = set_team_layout(b.get(), opts && opts->tiles_per_member > 0 ? opts->tiles_per_member
                                                                   : tuned_tiles_per_member(b.get()));
        if (rc) return rc;

        // uploads
        HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&b->ev0));
        HIP_TRY(hipEventCreate(&b->ev1));
        const size_t N = (size_t)P.n_total;
        std::vector<double2> x0(N);
        for (int k = 0; k < count; k++) {
            const nr::PlanNet &pn = P.nets[k];
            const nr::Network &net = *b->nets[k];
            for (int d = 0; d < pn.n_nodes; d++) {
                int o = pn.perm[d];
                x0[pn.node_base + d] = make_double2(net.pos[2 * o], net.pos[2 * o + 1]);
            }
        }
        std::vector<double2> zeros(N, make_double2(0.0, 0.0));
        HIP_TRY(b->x_init.upload(x0));
        if (b->precision == NR_PREC_MIXED) {
            std::vector<float2> z32(N, make_float2(0.0f, 0.0f));
            for (int s = 0; s < 2; s++) {
                HIP_TRY(b->u32[s].upload(z32));
                HIP_TRY(b->v32[s].upload(z32));
                HIP_TRY(b->f32[s].upload(z32));
            }
        }
        for (int s = 0; s < 2; s++) {
            HIP_TRY(b->x[s].upload(x0));
            HIP_TRY(b->v[s].upload(zeros));
            HIP_TRY(b->f[s].upload(zeros));
        }
        HIP_TRY(b->mask.upload(P.mask));
        {
            std::vector<uint4> blob16(P.blob.size() / 16);
            if (!blob16.empty()) std::memcpy(blob16.data(), P.blob.data(), P.blob.size());
            HIP_TRY(b->blob.upload(blob16));
        }
        HIP_TRY(b->tiles.upload(P.tiles));
        b->hnet.resize(count);
        for (int k = 0; k < count; k++) {
            nr::NetDev &d = b->hnet[k];
            std::memset(&d, 0, sizeof(d));
            const nr::PlanNet &pn = P.nets[k];
            const nr::Network &net = *b->nets[k];
            d.node_base = pn.node_base;
            d.n_nodes = pn.n_nodes;
            d.tile_base = pn.tile_base;
            d.n_tiles = pn.n_tiles;
            d.width = net.width;
            d.height = net.height;
            d.miny = net.miny;
            d.yrange = net.height > 0 ? net.height : net.maxy - net.miny;
            d.ks = b->ks[k];
            d.kb = b->kb[k];
        }
        HIP_TRY(b->dnet.upload(b->hnet));
        HIP_TRY(b->dout.alloc(count));
        HIP_TRY(b->run.alloc(count));
        HIP_TRY(b->err.alloc(1));
        HIP_TRY(hipMemset(b->err.p, 0, sizeof(unsigned)));
        *out = b.release();
        return NR_OK;
    });
}

int nr_batch_strain(nr_batch *b, int index, double xdisp, double ydisp)
{
    if (!b) return fail(NR_E_ARG, "null batch");
    int lo = index < 0 ? 0 : index, hi = index < 0 ? (int)b->hnet.size() : index + 1;
    if (index >= (int)b->hnet.size()) return fail(NR_E_ARG, "index out of range");
    for (int k = lo; k < hi; k++) {
        if (b->hnet[k].height > 0 && ydisp != 0.0)
            return fail(NR_E_ARG, "compression is not defined in the Lees-Edwards box");
        b->hnet[k].pend_x += xdisp;
        b->hnet[k].pend_y += ydisp;
    }
    return NR_OK;
}

static void fill_kargs(const nr_batch *b, int mode, nr::KArgs &a)
{
    std::memset(&a, 0, sizeof(a));
    for (int s = 0; s < 2; s++) {
        a.x[s] = b->x[s].p;
        a.v[s] = b->v[s].p;
        a.f[s] = b->f[s].p;
    }
    for (int s = 0; s < 2; s++) {
        a.u32[s] = b->u32[s].p;
        a.v32[s] = b->v32[s].p;
        a.f32[s] = b->f32[s].p;
    }
    a.mixed = b->precision == NR_PREC_MIXED ? 1 : 0;
    a.switch_ratio = b->switch_ratio;
    a.stall_steps = b->stall_steps;
    a.mask = b->mask.p;
    a.blob = b->blob.p;
    a.tiles = b->tiles.p;
    a.nets = b->dnet.p;
    a.out = b->dout.p;
    a.tile_cap = (int)(b->lds - nr::kCacheBytes);
    a.mode = mode;
    a.err = b->err.p;
    a.timeout_ticks = 100000000ll * 60;  // 60 s at the 100 MHz constant clock
    a.run = b->run.p;
    a.incs = b->incs.p;
    a.inc_out = b->inc_out.p;
    a.slice_steps = 0;
}

static int run_kernel(nr_batch *b, int mode)
{
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipMemcpyAsync(b->dnet.p, b->hnet.data(), sizeof(nr::NetDev) * b->hnet.size(),
                           hipMemcpyHostToDevice, b->stream));
    nr::KArgs a;
    fill_kargs(b, mode, a);
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    // every call starts fresh relaxes
    HIP_TRY(hipMemsetAsync(b->run.p, 0, sizeof(nr::NetRun) * b->run.n, b->stream));
    for (auto &ld : b->launches) {
        HIP_TRY(hipMemsetAsync(ld->queues.p, 0, sizeof(int) * ld->queues.n, b->stream));
        // partial-sum records carry per-launch step tags starting at 1
        HIP_TRY(hipMemsetAsync(ld->partials.p, 0, sizeof(unsigned long long) * ld->partials.n,
                               b->stream));
        a.roles = ld->roles.p;
        a.class_jobs = ld->class_jobs.p;
        a.class_off = ld->class_off.p;
        a.queues = ld->queues.p;
        a.partials = ld->partials.p;
        a.bar = ld->bar.p;
        a.mailbox = ld->mailbox.p;
        std::memcpy(a.cls_of_log2g, ld->L.cls_of_log2g, sizeof(a.cls_of_log2g));
        const int method = mode == nr::kModeEval ? NR_METHOD_FIRE : b->method;
        int e = nr::launch_relax(a, b->block, b->npt, method, ld->L.grid, b->lds, ld->L.cooperative,
                                 b->stream);
        if (e) return hip_fail((hipError_t)e, "relax kernel launch");
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    unsigned err = 0;
    HIP_TRY(hipMemcpy(&err, b->err.p, sizeof(unsigned), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(b->hnet.data(), b->dnet.p, sizeof(nr::NetDev) * b->hnet.size(),
                      hipMemcpyDeviceToHost));
    if (err) {
        (void)hipMemset(b->err.p, 0, sizeof(unsigned));
        return fail(NR_E_TIMEOUT, "a team barrier exceeded its wait bound");
    }
    return NR_OK;
}

static int set_params(nr_batch *b, const nr_relax_params *params, int shared)
{
    for (size_t k = 0; k < b->hnet.size(); k++)
        if ((shared ? params[0] : params[k]).method != params[0].method)
            return fail(NR_E_ARG, "one relax call uses one method for every network");
    b->method = params[0].method;
    for (size_t k = 0; k < b->hnet.size(); k++) {
        const nr_relax_params &p = shared ? params[0] : params[k];
        if (p.method != NR_METHOD_FIRE && p.method != NR_METHOD_CG && p.method != NR_METHOD_CG_PR)
            return fail(NR_E_ARG, "unknown relax method");
        if (!(p.mass > 0) || p.max_steps < 0 || p.fcut < 0)
            return fail(NR_E_ARG, "invalid relax parameters");
        nr::NetDev &d = b->hnet[k];
        d.a0 = p.alpha0;
        d.fa = p.f_alpha;
        d.dtmax = p.dt_max;
        d.finc = p.f_inc;
        d.fdec = p.f_dec;
        d.nmin = p.nmin;
        d.mass = p.mass;
        d.fcut = p.fcut;
        d.max_steps = p.max_steps;
        d.crit = p.criterion;
        d.method = p.method;
    }
    return NR_OK;
}

// Step counts of the batch's last relax become the cost estimates of the
// next one (the class allocation and queue order only: team sizes, and so
// the arithmetic, do not change).
static int plan_from_history(nr_batch *b)
{
    // a captured graph holds the current tables: keep them while it lives
    if (b->hist_steps.empty() || b->hist_planned || b->gexec) return NR_OK;
    b->hist_planned = true;
    return set_team_layout(b, b->tiles_per_member);
}

static void record_history(nr_batch *b, const std::vector<double> &steps)
{
    b->hist_steps = steps;
    for (double &v : b->hist_steps) v = std::max(v, 1.0);
    b->hist_planned = false;
}

int nr_batch_relax(nr_batch *b, const nr_relax_params *params, int shared, nr_report *reports)
{
    if (!b || !params) return fail(NR_E_ARG, "null argument");
    int rc = set_params(b, params, shared);
    if (rc) return rc;
    rc = plan_from_history(b);
    if (rc) return rc;
    rc = run_kernel(b, nr::kModeRelax);
    if (rc) return rc;
    std::vector<nr::NetOut> o(b->hnet.size());
    HIP_TRY(hipMemcpy(o.data(), b->dout.p, sizeof(nr::NetOut) * o.size(), hipMemcpyDeviceToHost));
    std::vector<double> steps(o.size());
    for (size_t k = 0; k < o.size(); k++) {
        steps[k] = (double)o[k].steps;
        if (reports) nr::to_report(o[k], reports[k]);
    }
    record_history(b, steps);
    return NR_OK;
}

int nr_batch_ramp(nr_batch *b, int n_inc, const double *incs, const nr_relax_params *params,
                  int shared, nr_report *reports)
{
    if (!b || !params || (n_inc > 0 && !incs) || n_inc < 0)
        return fail(NR_E_ARG, "invalid arguments");
    if (n_inc == 0) return NR_OK;
    const int count = (int)b->hnet.size();
    for (int i = 0; i < n_inc; i++) {
        if (!std::isfinite(incs[2 * i]) || !std::isfinite(incs[2 * i + 1]))
            return fail(NR_E_ARG, "non-finite strain increment");
        for (int k = 0; k < count; k++)
            if (b->hnet[k].height > 0 && incs[2 * i + 1] != 0.0)
                return fail(NR_E_ARG, "compression is not defined in the Lees-Edwards box");
    }
    int rc = set_params(b, params, shared);
    if (rc) return rc;
    std::vector<double2> h(n_inc);
    for (int i = 0; i < n_inc; i++) h[i] = make_double2(incs[2 * i], incs[2 * i + 1]);
    HIP_TRY(b->incs.upload(h));
    if (b->inc_out.n < (size_t)count * n_inc) HIP_TRY(b->inc_out.alloc((size_t)count * n_inc));
    for (int k = 0; k < count; k++) {
        b->hnet[k].inc_count = n_inc;
        b->hnet[k].inc_base = 0;
        b->hnet[k].out_base = k * n_inc;
    }
    rc = plan_from_history(b);
    if (!rc) rc = run_kernel(b, nr::kModeRelax);
    for (int k = 0; k < count; k++) b->hnet[k].inc_count = 0;
    if (rc) return rc;
    std::vector<nr::NetOut> o((size_t)count * n_inc);
    HIP_TRY(hipMemcpy(o.data(), b->inc_out.p, sizeof(nr::NetOut) * o.size(),
                      hipMemcpyDeviceToHost));
    std::vector<double> steps(count, 0.0);
    for (int k = 0; k < count; k++)
        for (int i = 0; i < n_inc; i++) {
            // a ramp that stopped early leaves its later reports untouched
            const nr::NetOut &r = o[(size_t)k * n_inc + i];
            if (r.status == NR_CONVERGED || r.status == NR_NOT_CONVERGED) steps[k] += r.steps;
        }
    if (reports)
        for (size_t i = 0; i < o.size(); i++) nr::to_report(o[i], reports[i]);
    record_history(b, steps);
    return NR_OK;
}

// Records one strain step of every network (pending step += its affine step,
// fresh relax state, the relax launches, copies of the reports and records
// into pinned host memory) on the batch stream.  Fill kernels stand in for
// memsets; team launches are ordinary launches (their grid never exceeds the
// resident workgroups, so the teams are co-resident on an otherwise idle GPU,
// and every team wait is bounded).
static int capture_step(nr_batch *b)
{
    const int count = (int)b->hnet.size();
    nr::KArgs a;
    fill_kargs(b, nr::kModeRelax, a);
    hipStream_t st = b->stream;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    int e = nr::launch_strain_step(b->dnet.p, count, b->gstep.p, st);
    if (!e) e = nr::launch_zero(b->run.p, sizeof(nr::NetRun) * b->run.n, st);
    for (auto &ld : b->launches) {
        if (!e) e = nr::launch_zero(ld->queues.p, sizeof(int) * ld->queues.n, st);
        if (!e) e = nr::launch_zero(ld->partials.p, sizeof(unsigned long long) * ld->partials.n, st);
        a.roles = ld->roles.p;
        a.class_jobs = ld->class_jobs.p;
        a.class_off = ld->class_off.p;
        a.queues = ld->queues.p;
        a.partials = ld->partials.p;
        a.bar = ld->bar.p;
        a.mailbox = ld->mailbox.p;
        std::memcpy(a.cls_of_log2g, ld->L.cls_of_log2g, sizeof(a.cls_of_log2g));
        if (!e) e = nr::launch_relax(a, b->block, b->npt, b->method, ld->L.grid, b->lds, false, st);
    }
    if (!e) e = (int)hipMemcpyAsync(b->h_out, b->dout.p, sizeof(nr::NetOut) * count,
                                    hipMemcpyDeviceToHost, st);
    if (!e) e = (int)hipMemcpyAsync(b->h_net, b->dnet.p, sizeof(nr::NetDev) * count,
                                    hipMemcpyDeviceToHost, st);
    if (!e) e = (int)hipMemcpyAsync(b->h_err, b->err.p, sizeof(unsigned), hipMemcpyDeviceToHost, st);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    if (e) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail((hipError_t)e, "strain step capture");
    }
    if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture");
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    return NR_OK;
}

int nr_batch_graph_capture(nr_batch *b, const double *xdisp, const double *ydisp,
                           const nr_relax_params *params, int shared)
{
    if (!b || !xdisp || !ydisp || !params) return fail(NR_E_ARG, "null argument");
    const int count = (int)b->hnet.size();
    std::vector<double2> d(count);
    for (int k = 0; k < count; k++) {
        if (!std::isfinite(xdisp[k]) || !std::isfinite(ydisp[k]))
            return fail(NR_E_ARG, "non-finite strain step");
        if (b->hnet[k].height > 0 && ydisp[k] != 0.0)
            return fail(NR_E_ARG, "compression is not defined in the Lees-Edwards box");
        d[k] = make_double2(xdisp[k], ydisp[k]);
    }
    int rc = set_params(b, params, shared);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(b->device));
    b->drop_graph();
    HIP_TRY(b->gstep.upload(d));
    if (!b->h_out) {
        HIP_TRY(hipHostMalloc((void **)&b->h_out, sizeof(nr::NetOut) * count, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc((void **)&b->h_net, sizeof(nr::NetDev) * count, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc((void **)&b->h_err, sizeof(unsigned), hipHostMallocDefault));
    }
    // pick the kernels' LDS size before capturing their launches
    int per_cu = 0, n_cu = 0;
    HIP_TRY((hipError_t)nr::max_resident_blocks(b->block, b->npt, b->lds, &per_cu, &n_cu));
    return capture_step(b);
}

int nr_batch_graph_replay(nr_batch *b, nr_report *reports)
{
    if (!b || !b->gexec) return fail(NR_E_ARG, "no captured strain step");
    HIP_TRY(hipSetDevice(b->device));
    // the host's records (positions set, steps added since) go first
    HIP_TRY(hipMemcpyAsync(b->dnet.p, b->hnet.data(), sizeof(nr::NetDev) * b->hnet.size(),
                           hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    HIP_TRY(hipGraphLaunch(b->gexec, b->stream));
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    std::memcpy(b->hnet.data(), b->h_net, sizeof(nr::NetDev) * b->hnet.size());
    if (*b->h_err) {
        (void)hipMemset(b->err.p, 0, sizeof(unsigned));
        return fail(NR_E_TIMEOUT, "a team barrier exceeded its wait bound");
    }
    if (reports)
        for (size_t k = 0; k < b->hnet.size(); k++) nr::to_report(b->h_out[k], reports[k]);
    return NR_OK;
}

static int copy_state(nr_batch *b, int index, const DevBuf<double2> *bufs, double *out)
{
    if (!b || !out || index < 0 || index >= (int)b->hnet.size())
        return fail(NR_E_ARG, "invalid arguments");
    const nr::PlanNet &pn = b->plan.nets[index];
    std::vector<double2> tmp(pn.n_nodes);
    const int par = b->hnet[index].parity;
    HIP_TRY(hipMemcpy(tmp.data(), bufs[par].p + pn.node_base, sizeof(double2) * pn.n_nodes,
                      hipMemcpyDeviceToHost));
    for (int d = 0; d < pn.n_nodes; d++) {
        int o = pn.perm[d];
        out[2 * o] = tmp[d].x;
        out[2 * o + 1] = tmp[d].y;
    }
    return NR_OK;
}

int nr_batch_reset(nr_batch *b)
{
    if (!b) return fail(NR_E_ARG, "null batch");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipMemcpyAsync(b->x[0].p, b->x_init.p, sizeof(double2) * b->x_init.n,
                           hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (nr::NetDev &d : b->hnet) {
        d.parity = 0;
        d.le_shift = 0.0;
        d.pend_x = 0.0;
        d.pend_y = 0.0;
    }
    return NR_OK;
}

int nr_batch_get_positions(nr_batch *b, int index, double *pos)
{
    return copy_state(b, index, b ? b->x : nullptr, pos);
}

int nr_batch_set_positions(nr_batch *b, int index, const double *pos)
{
    if (!b || !pos || index < 0 || index >= (int)b->hnet.size())
        return fail(NR_E_ARG, "invalid arguments");
    const nr::PlanNet &pn = b->plan.nets[index];
    std::vector<double2> tmp(pn.n_nodes);
    for (int d = 0; d < pn.n_nodes; d++) {
        int o = pn.perm[d];
        tmp[d] = make_double2(pos[2 * o], pos[2 * o + 1]);
    }
    const int par = b->hnet[index].parity;
    HIP_TRY(hipMemcpy(b->x[par].p + pn.node_base, tmp.data(), sizeof(double2) * pn.n_nodes,
                      hipMemcpyHostToDevice));
    return NR_OK;
}

// Workgroup partials (non-affinity, then axial energy, strain sum, max
// |strain| per workgroup) added in workgroup order; the bond kernel does not
// run without bonds, so its partials are not read then.
static void analysis_result(const double *hp, int n_nodes, int ne, double k_axial,
                            nr_analysis *out)
{
    double na = 0.0, ax = 0.0, sum = 0.0, mx = 0.0;
    for (int w = 0; w < nr::kAnalysisBlocks; w++) {
        na += hp[w];
        if (ne > 0) {
            ax += hp[nr::kAnalysisBlocks + 3 * w];
            sum += hp[nr::kAnalysisBlocks + 3 * w + 1];
            mx = std::max(mx, hp[nr::kAnalysisBlocks + 3 * w + 2]);
        }
    }
    out->nonaffinity = n_nodes > 0 ? na / n_nodes : 0.0;
    out->axial_energy = k_axial * ax;
    out->mean_bond_strain = ne > 0 ? sum / ne : 0.0;
    out->max_bond_strain = mx;
}

int nr_batch_analyze(nr_batch *b, int index, double shear_strain, double axial_strain,
                     double k_axial, nr_analysis *out, double *bond_strains)
{
    if (!b || !out || index < 0 || index >= (int)b->hnet.size())
        return fail(NR_E_ARG, "invalid arguments");
    HIP_TRY(hipSetDevice(b->device));
    const nr::Network &net = *b->nets[index];
    const nr::PlanNet &pn = b->plan.nets[index];
    const nr::NetDev &nd = b->hnet[index];
    const int ne = net.e();
    if (b->an_bonds.size() != b->hnet.size()) b->an_bonds.resize(b->hnet.size());
    if (!b->an_bonds[index] && ne > 0) {
        std::vector<int4> h(ne);
        for (int e = 0; e < ne; e++)
            h[e] = make_int4(pn.iperm[net.bonds[2 * e]], pn.iperm[net.bonds[2 * e + 1]],
                             net.bwrap[2 * e], net.bwrap[2 * e + 1]);
        auto buf = std::make_unique<DevBuf<int4>>();
        HIP_TRY(buf->upload(h));
        b->an_bonds[index] = std::move(buf);
    }
    DevBuf<double> strain, part;
    HIP_TRY(strain.alloc(std::max(ne, 1)));
    HIP_TRY(part.alloc(4 * nr::kAnalysisBlocks));
    const double2 *x = b->x[nd.parity].p + pn.node_base;
    const double2 *x0 = b->x_init.p + pn.node_base;
    HIP_TRY((hipError_t)nr::launch_analysis(
        x, x0, pn.n_nodes, net.miny, shear_strain, axial_strain,
        ne > 0 ? b->an_bonds[index]->p : nullptr, ne, net.width, net.height, nd.le_shift,
        strain.p, part.p, b->stream));
    std::vector<double> hp(4 * nr::kAnalysisBlocks);
    HIP_TRY(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                           b->stream));
    if (bond_strains && ne > 0)
        HIP_TRY(hipMemcpyAsync(bond_strains, strain.p, sizeof(double) * ne,
                               hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    analysis_result(hp.data(), pn.n_nodes, ne, k_axial, out);
    return NR_OK;
}

// nr_batch_analyze with its workspace in the library's fixed device array
// (analysis.hip): no device or host allocation.  The bond list comes from an
// earlier nr_batch_analyze call when there was one, else it is staged into
// the workspace from the network in chunks through a stack buffer.
int nr_batch_analyze_noalloc(nr_batch *b, int index, double shear_strain, double axial_strain,
                             double k_axial, nr_analysis *out, double *bond_strains)
{
    if (!b || !out || index < 0 || index >= (int)b->hnet.size())
        return fail(NR_E_ARG, "invalid arguments");
    const nr::Network &net = *b->nets[index];
    const nr::PlanNet &pn = b->plan.nets[index];
    const nr::NetDev &nd = b->hnet[index];
    const int ne = net.e();
    if (ne > nr::kNoAllocMaxBonds)
        return fail(NR_E_ARG, "nr_batch_analyze_noalloc: " + std::to_string(ne) +
                                  " bonds exceed its workspace of " +
                                  std::to_string(nr::kNoAllocMaxBonds));
    HIP_TRY(hipSetDevice(b->device));
    double *ws = nullptr;
    HIP_TRY((hipError_t)nr::analysis_scratch(b->device, &ws));
    double *part = ws, *strain = ws + 4 * nr::kAnalysisBlocks;
    int4 *staged = reinterpret_cast<int4 *>(strain + nr::kNoAllocMaxBonds);
    const int4 *bonds = nullptr;
    if (ne > 0) {
        if (b->an_bonds.size() == b->hnet.size() && b->an_bonds[index]) {
            bonds = b->an_bonds[index]->p;
        } else {
            constexpr int kChunk = 1024;
            int4 chunk[kChunk];
            for (int e0 = 0; e0 < ne; e0 += kChunk) {
                const int m = std::min(kChunk, ne - e0);
                for (int i = 0; i < m; i++) {
                    const int e = e0 + i;
                    chunk[i] = make_int4(pn.iperm[net.bonds[2 * e]], pn.iperm[net.bonds[2 * e + 1]],
                                         net.bwrap[2 * e], net.bwrap[2 * e + 1]);
                }
                HIP_TRY(hipMemcpyAsync(staged + e0, chunk, sizeof(int4) * m,
                                       hipMemcpyHostToDevice, b->stream));
                // the stack chunk is reused: wait for its copy
                HIP_TRY(hipStreamSynchronize(b->stream));
            }
            bonds = staged;
        }
    }
    const double2 *x = b->x[nd.parity].p + pn.node_base;
    const double2 *x0 = b->x_init.p + pn.node_base;
    HIP_TRY((hipError_t)nr::launch_analysis(x, x0, pn.n_nodes, net.miny, shear_strain,
                                            axial_strain, bonds, ne, net.width, net.height,
                                            nd.le_shift, bond_strains ? strain : nullptr, part,
                                            b->stream));
    double hp[4 * nr::kAnalysisBlocks];
    HIP_TRY(hipMemcpyAsync(hp, part, sizeof(hp), hipMemcpyDeviceToHost, b->stream));
    if (bond_strains && ne > 0)
        HIP_TRY(hipMemcpyAsync(bond_strains, strain, sizeof(double) * ne,
                               hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    analysis_result(hp, pn.n_nodes, ne, k_axial, out);
    return NR_OK;
}

int nr_batch_eval(nr_batch *b, int index, double *force, double *energies)
{
    if (!b || index < 0 || index >= (int)b->hnet.size()) return fail(NR_E_ARG, "invalid arguments");
    int rc = run_kernel(b, nr::kModeEval);
    if (rc) return rc;
    if (force) {
        rc = copy_state(b, index, b->f, force);
        if (rc) return rc;
    }
    if (energies) {
        nr::NetOut o;
        HIP_TRY(hipMemcpy(&o, b->dout.p + index, sizeof(o), hipMemcpyDeviceToHost));
        energies[0] = o.es;
        energies[1] = o.eb;
    }
    return NR_OK;
}

double nr_batch_last_ms(const nr_batch *b) { return b ? b->last_ms : 0.0; }

int nr_batch_counts(const nr_batch *b, int index, int64_t *nodes, int64_t *bonds,
                    int64_t *triples)
{
    if (!b) return fail(NR_E_ARG, "null batch");
    int64_t n = 0, e = 0, t = 0;
    for (size_t k = 0; k < b->plan.nets.size(); k++) {
        if (index >= 0 && (int)k != index) continue;
        n += b->plan.nets[k].n_nodes;
        e += b->plan.nets[k].n_bonds;
        t += b->plan.nets[k].n_triples;
    }
    if (nodes) *nodes = n;
    if (bonds) *bonds = e;
    if (triples) *triples = t;
    return NR_OK;
}

void nr_batch_destroy(nr_batch *b) { delete b; }

// Shape tuner.  Every candidate workgroup/tile shape and tiles-per-member
// rule ({1, 2, 4, 8}, or the caller's) is measured by a probe relax in which
// network k takes a step budget proportional to its expected step count (a
// fit over size and bond density; the budgets average probe_steps), teams
// planned from those budgets as a relax planned from known step counts is,
// and no convergence test.  The fastest candidate wins; its probe time scaled
// by the expected steps over the budgets is the prediction.  (Extrapolating
// a uniform probe with the makespan model ranked large teams too well: at 16
// tiles per member it predicted 6.7 s for a batch that took 10.3 s.)
int nr_autotune(int count, nr_network *const *nets, const double *ks, const double *kb,
                int probe_steps, nr_batch_opts *opts, double *predicted_ms)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (count <= 0 || !nets || !ks || !kb || !opts || probe_steps < 1)
            return fail(NR_E_ARG, "invalid arguments");
        struct Shape {
            int block, tile;
        };
        std::vector<Shape> shapes = {{256, 256}, {512, 512}};
        if (opts->precision == NR_PREC_FP64) shapes.push_back({256, 512});
        std::vector<int> cands = {1, 2, 4, kDefaultTilesPerMember};
        if (opts->tiles_per_member > 0) cands = {opts->tiles_per_member};
        nr_relax_params base;
        std::memset(&base, 0, sizeof(base));
        base.alpha0 = 0.1;
        base.f_alpha = 0.99;
        base.dt_max = 0.1;
        base.f_inc = 1.1;
        base.f_dec = 0.5;
        base.nmin = 5;
        base.mass = 1.0;
        base.fcut = 0.0;  // never met: every network takes its whole budget
        base.criterion = NR_CRIT_MAX;
        base.method = NR_METHOD_FIRE;
        double best = std::numeric_limits<double>::infinity();
        nr_batch_opts best_opts = *opts;
        for (const Shape &sh : shapes) {
            nr_batch_opts o = *opts;
            o.block = sh.block;
            o.tile_nodes = sh.tile;
            o.tiles_per_member = cands[0];
            nr_batch *raw = nullptr;
            int rc = nr_batch_create(count, nets, ks, kb, &o, &raw);
            if (rc == NR_E_TOPO || rc == NR_E_ARG) continue;  // shape cannot hold these networks
            if (rc) return rc;
            std::unique_ptr<nr_batch> b(raw);
            std::vector<double> full(count), budget(count);
            double sum_full = 0.0;
            for (int k = 0; k < count; k++) {
                full[k] = nr::expected_steps(b->plan.nets[k].n_nodes, b->plan.nets[k].n_bonds);
                sum_full += full[k];
            }
            const double frac = probe_steps * (double)count / sum_full;
            double sum_budget = 0.0;
            std::vector<nr_relax_params> prm(count, base);
            for (int k = 0; k < count; k++) {
                budget[k] = std::max(1.0, std::round(full[k] * frac));
                prm[k].max_steps = (int)budget[k];
                sum_budget += budget[k];
            }
            for (int tpm : cands) {
                b->hist_steps = budget;
                rc = set_team_layout(b.get(), tpm);
                if (rc) return rc;
                b->hist_planned = true;
                rc = nr_batch_relax(b.get(), prm.data(), 0, nullptr);
                if (rc) return rc;
                const double t = b->last_ms * sum_full / sum_budget;
                if (t < best) {
                    best = t;
                    best_opts = o;
                    best_opts.tiles_per_member = tpm;
                }
            }
        }
        if (!std::isfinite(best)) return fail(NR_E_TOPO, "no shape holds these networks");
        *opts = best_opts;
        if (predicted_ms) *predicted_ms = best;
        return NR_OK;
    });
}

}  // extern "C"
