// Network + FEM-gel composite on the device: the C-ABI of nr_mesh /
// nr_composite and the batch protocol of the reference's composite program
// (src/network_mesh_relax.cu: relaxation_batch :1393-1421, do_relaxation_run
// :1095-1391, run_fire_md :865-1066).
//
// A relax uploads the FIRE scalars, zeroes v, computes the first force and
// drift, then replays a captured chunk of kMeshChunk steps (three kernels
// each, mesh_kernels.hip) until the device's done flag is up; positions stay
// resident between the increments of a ramp.  With NRELAX_GRAPHS=0 in the
// environment at nr_composite_create, each chunk's kernels are launched
// directly instead (same kernels, same order; for tools that do not follow
// graph launches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nrelax.h"
#include "internal.hpp"
#include "mesh.hpp"
#include "mesh_device.hpp"
#include "protocol.hpp"

struct nr_mesh {
    nr::Mesh m;
};

namespace {

constexpr int kMeshChunk = 32;  // FIRE steps per captured graph (even)

}  // namespace

struct nr_composite {
    nr::Mesh mesh;
    nr::MeshSetup setup;
    nr_gel_params gel{};
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    nr::DevBuf<double> x, v, f, part, tot, state, epart, inv, area;
    nr::DevBuf<unsigned> ticket;
    nr::DevBuf<double2> slot, bond_geo, tri_off;
    nr::DevBuf<int2> bond;
    nr::DevBuf<int4> tri, facet;
    nr::DevBuf<int> ent, cnt, twin;
    nr::DevBuf<unsigned char> fcode;
    nr::MeshDev md{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    nr_relax_params captured{};
    bool have_graph = false;
    bool graphs = true;  // false: each chunk's kernels launched directly
    int energy_blocks = 1;
    double last_ms = 0.0;

    ~nr_composite()
    {
        // a relax that returned early (an error, the step bound) may have
        // left graph replays queued: they finish before the graph goes
        if (st) (void)hipStreamSynchronize(st);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace nr {
namespace {

// Slot entries per point, (slot << 1) | negate, in element order, in three
// sections: bonds (bond b is slot b, + on its first point, - on its second),
// triples (triple t is slots nb + t for end a and nb + nt + t for end b,
// both negated on the hinge) and facets (facet f is slot
// nb + 2nt + k nf + f on its corner k).  Each section is column-major with the section's largest
// per-point count as its width: entry k of point i at sec[s] + k * n + i.
// cnt packs the three counts of a point in 8 bits each.
int build_gather(int n, const std::vector<int> &bonds, int nb, const std::vector<int> &tri, int nt,
                 const std::vector<int> &facets, int nf, std::vector<int> &ent,
                 std::vector<int> &cnt, long sec[3])
{
    std::vector<std::vector<int>> lists(3 * (size_t)n);
    auto put = [&](int s, int point, int e) { lists[3 * (size_t)point + s].push_back(e); };
    for (int b = 0; b < nb; b++) {
        put(0, bonds[2 * b], b << 1);
        put(0, bonds[2 * b + 1], (b << 1) | 1);
    }
    for (int t = 0; t < nt; t++) {
        const int sa = nb + t, sb = nb + nt + t;
        put(1, tri[3 * t], (sa << 1) | 1);
        put(1, tri[3 * t], (sb << 1) | 1);
        put(1, tri[3 * t + 1], sa << 1);
        put(1, tri[3 * t + 2], sb << 1);
    }
    for (int f = 0; f < nf; f++)
        for (int k = 0; k < 3; k++) put(2, facets[3 * f + k], (nb + 2 * nt + k * nf + f) << 1);
    size_t width[3] = {0, 0, 0};
    for (int i = 0; i < n; i++)
        for (int s = 0; s < 3; s++)
            width[s] = std::max(width[s], lists[3 * (size_t)i + s].size());
    if (width[0] > 255 || width[1] > 255 || width[2] > 255)
        return fail(NR_E_TOPO, "a point has more than 255 entries of one element type");
    sec[0] = 0;
    sec[1] = (long)(width[0] * (size_t)n);
    sec[2] = sec[1] + (long)(width[1] * (size_t)n);
    ent.assign((size_t)sec[2] + width[2] * (size_t)n + 1, 0);
    cnt.assign((size_t)n, 0);
    for (int i = 0; i < n; i++) {
        for (int s = 0; s < 3; s++) {
            const std::vector<int> &l = lists[3 * (size_t)i + s];
            for (size_t k = 0; k < l.size(); k++) ent[(size_t)sec[s] + k * (size_t)n + i] = l[k];
        }
        cnt[i] = (int)lists[3 * (size_t)i].size() | (int)lists[3 * (size_t)i + 1].size() << 8 |
                 (int)lists[3 * (size_t)i + 2].size() << 16;
    }
    return NR_OK;
}

int composite_upload(nr_composite &c)
{
    const Network &net = c.mesh.net;
    const MeshSetup &s = c.setup;
    const int n = net.n;
    // the reference skips a term whose modulus is 0 (:928-951)
    const int nb = c.gel.ks > 0 ? net.e() : 0;
    const int nt = c.gel.kb > 0 ? net.t() : 0;
    const int nf = c.gel.bmod > 0 ? c.mesh.nf() : 0;
    std::vector<int> ent, cnt;
    long sec[3];
    int rc = build_gather(n, net.bonds, nb, net.tri, nt, c.mesh.facets, nf, ent, cnt, sec);
    if (rc) return rc;
    std::vector<int2> bond((size_t)nb);
    std::vector<double2> geo((size_t)nb);
    for (int b = 0; b < nb; b++) {
        bond[b] = make_int2(net.bonds[2 * b], net.bonds[2 * b + 1]);
        geo[b] = make_double2(net.rest[b], s.width * net.bwrap[2 * b]);
    }
    std::vector<int4> tri((size_t)nt);
    std::vector<double2> toff((size_t)nt);
    for (int t = 0; t < nt; t++) {
        tri[t] = make_int4(net.tri[3 * t], net.tri[3 * t + 1], net.tri[3 * t + 2], 0);
        toff[t] = make_double2(s.width * net.twrap[4 * t], s.width * net.twrap[4 * t + 2]);
    }
    std::vector<int4> fac((size_t)nf);
    for (int f = 0; f < nf; f++)
        fac[f] = make_int4(c.mesh.facets[3 * f], c.mesh.facets[3 * f + 1], c.mesh.facets[3 * f + 2],
                           0);
    std::vector<double> inv(s.inv.begin(), s.inv.begin() + 6 * (size_t)nf);
    std::vector<double> area(c.mesh.areas.begin(), c.mesh.areas.begin() + nf);
    std::vector<int> twin;
    if (nf > 0 && c.mesh.np() > 0) {  // duplicate_forces runs with the gel only (:946-951)
        twin.assign((size_t)n, -1);
        for (int p = 0; p < c.mesh.np(); p++) {
            twin[(size_t)c.mesh.twins[2 * p]] = c.mesh.twins[2 * p + 1];
            twin[(size_t)c.mesh.twins[2 * p + 1]] = c.mesh.twins[2 * p];
        }
    }
    const size_t nslot = (size_t)nb + 2 * (size_t)nt + 3 * (size_t)nf;
    if (nslot >= (1u << 30)) return fail(NR_E_TOPO, "composite too large for the slot index");
    HIP_TRY(c.x.upload(s.pos));
    HIP_TRY(c.v.alloc(2 * (size_t)n));
    HIP_TRY(c.f.alloc(2 * (size_t)n));
    HIP_TRY(c.bond.upload(bond));
    HIP_TRY(c.bond_geo.upload(geo));
    HIP_TRY(c.tri.upload(tri));
    HIP_TRY(c.tri_off.upload(toff));
    HIP_TRY(c.facet.upload(fac));
    HIP_TRY(c.inv.upload(inv));
    HIP_TRY(c.area.upload(area));
    HIP_TRY(c.slot.alloc(nslot));
    HIP_TRY(c.ent.upload(ent));
    HIP_TRY(c.cnt.upload(cnt));
    if (!twin.empty()) HIP_TRY(c.twin.upload(twin));
    HIP_TRY(c.fcode.upload(net.fcode));
    HIP_TRY(c.part.alloc(4 * (size_t)kMeshMaxBlocks));
    HIP_TRY(c.tot.alloc(4));
    HIP_TRY(c.ticket.alloc(1));
    HIP_TRY(hipMemset(c.ticket.p, 0, sizeof(unsigned)));
    HIP_TRY(c.state.alloc(kStWords));
    HIP_TRY(c.epart.alloc(3 * (size_t)kMeshMaxBlocks));

    MeshDev &m = c.md;
    m = MeshDev{};
    m.n = n;
    m.nb = nb;
    m.nt = nt;
    m.nf = nf;
    m.x = c.x.p;
    m.v = c.v.p;
    m.f = c.f.p;
    m.bond = c.bond.p;
    m.bond_geo = c.bond_geo.p;
    m.tri = c.tri.p;
    m.tri_off = c.tri_off.p;
    m.facet = c.facet.p;
    m.inv = c.inv.p;
    m.area = c.area.p;
    m.slot = c.slot.p;
    m.ent = c.ent.p;
    m.cnt = c.cnt.p;
    for (int k = 0; k < 3; k++) m.sec[k] = sec[k];
    m.twin = twin.empty() ? nullptr : c.twin.p;
    m.fcode = c.fcode.p;
    m.part = c.part.p;
    m.tot = c.tot.p;
    m.ticket = c.ticket.p;
    m.st = c.state.p;
    m.gather_blocks = mesh_point_blocks(n);
    m.ks = c.gel.ks;
    m.kb = c.gel.kb;
    m.smod = c.gel.smod;
    m.bmod = c.gel.bmod;
    m.pref_d = s.pref_d;
    m.sqrt_dof = std::sqrt((double)(2 * n));
    c.energy_blocks = mesh_element_blocks(m);
    return NR_OK;
}

void set_fire(MeshDev &m, const nr_relax_params &p)
{
    m.alpha0 = p.alpha0;
    m.f_alpha = p.f_alpha;
    m.dt_max = p.dt_max;
    m.f_inc = p.f_inc;
    m.f_dec = p.f_dec;
    m.mass = p.mass;
    m.fcut = p.fcut;
    m.nmin = p.nmin;
    m.max_steps = p.max_steps;
    m.criterion = p.criterion;
}

bool same_params(const nr_relax_params &a, const nr_relax_params &b)
{
    return a.alpha0 == b.alpha0 && a.f_alpha == b.f_alpha && a.dt_max == b.dt_max &&
           a.f_inc == b.f_inc && a.f_dec == b.f_dec && a.nmin == b.nmin && a.mass == b.mass &&
           a.fcut == b.fcut && a.max_steps == b.max_steps && a.criterion == b.criterion;
}

// The FIRE scalars of a fresh relax: slot 0 = (dt_max, alpha0, 0, not done,
// step 0), slot 1 and the result cleared.
hipError_t reset_state(nr_composite &c, const nr_relax_params &p)
{
    double h[kStWords] = {0};
    h[kStDt] = p.dt_max;
    h[kStAlpha] = p.alpha0;
    hipError_t e = hipMemcpyAsync(c.state.p, h, sizeof(h), hipMemcpyHostToDevice, c.st);
    if (e == hipSuccess) e = hipStreamSynchronize(c.st);
    return e;
}

hipError_t capture_chunk(nr_composite &c)
{
    // nothing may still run the graph being replaced (a relax that returned
    // early can leave replays queued on the stream)
    hipError_t e = hipStreamSynchronize(c.st);
    if (e != hipSuccess) return e;
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.exec = nullptr;
    c.graph = nullptr;
    c.have_graph = false;
    e = hipStreamBeginCapture(c.st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    // the chunk starts reading slot 1 (the first drift writes it)
    for (int k = 0; k < kMeshChunk && e == hipSuccess; k++)
        e = (hipError_t)launch_mesh_step(c.md, (k + 1) & 1, 0, c.st);
    hipError_t e2 = hipStreamEndCapture(c.st, &c.graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0);
    if (e == hipSuccess) c.have_graph = true;
    return e;
}

int energies(nr_composite &c, double out[3])
{
    std::vector<double> h(3 * (size_t)c.energy_blocks);
    HIP_TRY((hipError_t)launch_mesh_energy(c.md, c.epart.p, c.energy_blocks, c.st));
    HIP_TRY(hipMemcpyAsync(h.data(), c.epart.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                           c.st));
    HIP_TRY(hipStreamSynchronize(c.st));
    out[0] = out[1] = out[2] = 0.0;
    for (int b = 0; b < c.energy_blocks; b++)
        for (int k = 0; k < 3; k++) out[k] += h[3 * (size_t)b + k];
    return NR_OK;
}

int relax(nr_composite &c, const nr_relax_params &p, nr_report &rep, double *e_gel)
{
    if (p.method != NR_METHOD_FIRE)
        return fail(NR_E_ARG, "composite relax supports FIRE only (the reference's method)");
    if (!(p.mass > 0) || !(p.dt_max > 0) || p.nmin < 0 || p.max_steps < 0)
        return fail(NR_E_ARG, "invalid FIRE parameters");
    HIP_TRY(hipSetDevice(c.device));
    set_fire(c.md, p);
    if (c.graphs && (!c.have_graph || !same_params(c.captured, p))) {
        HIP_TRY(capture_chunk(c));
        c.captured = p;
    }
    HIP_TRY(reset_state(c, p));
    HIP_TRY(hipMemsetAsync(c.v.p, 0, sizeof(double) * 2 * (size_t)c.md.n, c.st));
    HIP_TRY(hipEventRecord(c.e0, c.st));
    HIP_TRY((hipError_t)launch_mesh_step(c.md, 0, 1, c.st));
    double h[kStWords];
    const long max_replays = (long)p.max_steps / kMeshChunk + 3;
    long replays = 0;
    for (;;) {
        HIP_TRY(hipMemcpyAsync(h, c.state.p, sizeof(h), hipMemcpyDeviceToHost, c.st));
        HIP_TRY(hipStreamSynchronize(c.st));
        if (h[kStSlot + kStDone] != 0.0) break;
        if (replays++ > max_replays)
            return fail(NR_E_TIMEOUT, "composite relax did not finish within its step bound");
        if (c.graphs) {
            HIP_TRY(hipGraphLaunch(c.exec, c.st));
        } else {
            for (int k = 0; k < kMeshChunk; k++)
                HIP_TRY((hipError_t)launch_mesh_step(c.md, (k + 1) & 1, 0, c.st));
        }
    }
    HIP_TRY(hipEventRecord(c.e1, c.st));
    HIP_TRY(hipEventSynchronize(c.e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c.e0, c.e1));
    c.last_ms = ms;
    const double *r = h + 2 * kStSlot;
    double e[3];
    int rc = energies(c, e);
    if (rc) return rc;
    std::memset(&rep, 0, sizeof(rep));
    rep.rms_force = r[kResRms];
    rep.max_force = r[kResFmax];
    rep.e_stretch = e[0];
    rep.e_bend = e[1];
    rep.steps = (int)r[kResSteps];
    rep.status = (int)r[kResStatus];
    rep.fp64_steps = rep.steps;
    if (e_gel) *e_gel = e[2];
    return NR_OK;
}

}  // namespace
}  // namespace nr

// ---- C ABI ----------------------------------------------------------------

extern "C" int nr_mesh_read(const char *path, nr_mesh **out)
{
    return nr::guarded(NR_E_OPEN, [&]() -> int {
        if (!path || !out) return nr::fail(NR_E_ARG, "null argument");
        *out = nullptr;
        auto m = std::make_unique<nr_mesh>();
        std::string err;
        int rc = nr::read_mesh(path, m->m, err);
        if (rc) return nr::fail(rc, err);
        *out = m.release();
        return NR_OK;
    });
}

extern "C" int nr_mesh_lattice(int L, double p, uint64_t seed, nr_mesh **out)
{
    if (!out) return nr::fail(NR_E_ARG, "null argument");
    *out = nullptr;
    auto m = std::make_unique<nr_mesh>();
    std::string err;
    int rc = nr::make_mesh_lattice(L, p, seed, m->m, err);
    if (rc) return nr::fail(rc, err);
    *out = m.release();
    return NR_OK;
}

extern "C" int nr_mesh_write_nmc(const nr_mesh *m, const char *path)
{
    if (!m || !path) return nr::fail(NR_E_ARG, "null argument");
    int rc = nr::write_nmc(m->m, path);
    return rc ? nr::fail(rc, std::string("cannot write ") + path) : NR_OK;
}

extern "C" int nr_mesh_get_info(const nr_mesh *m, nr_mesh_info *info)
{
    if (!m || !info) return nr::fail(NR_E_ARG, "null argument");
    info->n_points = m->m.net.n;
    info->n_bonds = m->m.net.e();
    info->n_facets = m->m.nf();
    info->n_twins = m->m.np();
    info->offset = m->m.offset;
    info->minx = m->m.minx;
    info->maxx = m->m.maxx;
    info->miny = m->m.net.miny;
    info->maxy = m->m.net.maxy;
    return NR_OK;
}

extern "C" int nr_mesh_get_arrays(const nr_mesh *m, double *pos, int *facets, double *areas,
                                  int *twins)
{
    if (!m) return nr::fail(NR_E_ARG, "null argument");
    const nr::Mesh &s = m->m;
    if (pos) std::memcpy(pos, s.net.pos.data(), sizeof(double) * s.net.pos.size());
    if (facets) std::memcpy(facets, s.facets.data(), sizeof(int) * s.facets.size());
    if (areas) std::memcpy(areas, s.areas.data(), sizeof(double) * s.areas.size());
    if (twins) std::memcpy(twins, s.twins.data(), sizeof(int) * s.twins.size());
    return NR_OK;
}

extern "C" void nr_mesh_free(nr_mesh *m) { delete m; }

extern "C" int nr_composite_create(const nr_mesh *m, const nr_gel_params *g, int device,
                                   nr_composite **out)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!m || !g || !out) return nr::fail(NR_E_ARG, "null argument");
        *out = nullptr;
        auto c = std::make_unique<nr_composite>();
        c->mesh = m->m;
        c->gel = *g;
        c->device = device;
        const char *g_env = std::getenv("NRELAX_GRAPHS");
        c->graphs = !(g_env && g_env[0] == '0');
        std::string err;
        int rc = nr::mesh_setup(c->mesh, g->ks, g->kb, g->bmod, g->dilation, c->setup, err);
        if (rc) return nr::fail(rc, err);
        if (!(g->smod >= 0)) return nr::fail(NR_E_ARG, "smod must be >= 0");
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&c->e0));
        HIP_TRY(hipEventCreate(&c->e1));
        rc = nr::composite_upload(*c);
        if (rc) return rc;
        *out = c.release();
        return NR_OK;
    });
}

extern "C" int nr_composite_get_info(const nr_composite *c, nr_composite_info *info)
{
    if (!c || !info) return nr::fail(NR_E_ARG, "null argument");
    info->width = c->setup.width;
    info->range = c->setup.range;
    info->pref_d = c->setup.pref_d;
    info->net_bmod = c->setup.net_bmod;
    info->n_triples = c->md.nt;
    info->n_bonds = c->md.nb;
    info->n_facets = c->md.nf;
    info->n_points = c->md.n;
    return NR_OK;
}

extern "C" int nr_composite_affine(nr_composite *c, double yrange, double xdisp, double ydisp)
{
    if (!c) return nr::fail(NR_E_ARG, "null argument");
    if (!(yrange != 0.0) || !std::isfinite(xdisp) || !std::isfinite(ydisp))
        return nr::fail(NR_E_ARG, "invalid affine step");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY((hipError_t)nr::launch_mesh_affine(c->x.p, c->md.n, c->mesh.net.miny, yrange, xdisp,
                                               ydisp, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return NR_OK;
}

extern "C" int nr_composite_relax(nr_composite *c, const nr_relax_params *p, nr_report *rep,
                                  double *e_gel)
{
    if (!c || !p || !rep) return nr::fail(NR_E_ARG, "null argument");
    return nr::relax(*c, *p, *rep, e_gel);
}

extern "C" int nr_composite_get_positions(nr_composite *c, double *pos)
{
    if (!c || !pos) return nr::fail(NR_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(pos, c->x.p, sizeof(double) * 2 * (size_t)c->md.n,
                           hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return NR_OK;
}

extern "C" int nr_composite_set_positions(nr_composite *c, const double *pos)
{
    if (!c || !pos) return nr::fail(NR_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(c->x.p, pos, sizeof(double) * 2 * (size_t)c->md.n,
                           hipMemcpyHostToDevice, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return NR_OK;
}

// Force (clamped, twins summed) and {E_stretch, E_bend, E_gel} at the
// current positions, by the relax's own kernels.
extern "C" int nr_composite_eval(nr_composite *c, double *force, double *energies)
{
    if (!c) return nr::fail(NR_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    nr_relax_params p{};
    p.dt_max = 0.1;
    p.alpha0 = 0.1;
    p.mass = 1.0;
    p.max_steps = 0;  // the first-drift launch then stops at once: x stays
    nr::set_fire(c->md, p);
    HIP_TRY(nr::reset_state(*c, p));
    HIP_TRY(hipMemsetAsync(c->v.p, 0, sizeof(double) * 2 * (size_t)c->md.n, c->st));
    HIP_TRY((hipError_t)nr::launch_mesh_step(c->md, 0, 1, c->st));
    if (force)
        HIP_TRY(hipMemcpyAsync(force, c->f.p, sizeof(double) * 2 * (size_t)c->md.n,
                               hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (energies) {
        int rc = nr::energies(*c, energies);
        if (rc) return rc;
    }
    return NR_OK;
}

extern "C" double nr_composite_last_ms(const nr_composite *c) { return c ? c->last_ms : 0.0; }

extern "C" void nr_composite_destroy(nr_composite *c)
{
    if (c) (void)hipSetDevice(c->device);
    delete c;
}

// ---- the composite program's batch protocol ------------------------------

namespace nr {
namespace {

struct MeshBatchSpec {
    BatchSpec b;  // lines 1 and 3-6 as the network program's
    nr_gel_params gel{};
};

// parse_batch_file :413-500: as the network program's, except line 2 is
// "ks kb bmod [smod [dilation]]" (at least three numbers, none negative).
int parse_mesh_batch(const std::string &path, MeshBatchSpec &s, std::string &err)
{
    int rc = parse_batch_file(path, s.b, err);
    if (rc) return rc;
    std::ifstream in(path);
    std::string line;
    int got = 0;
    while (got < 2 && std::getline(in, line))
        if (!line.empty() && line[0] != '#') got++;
    nr_gel_params &g = s.gel;
    g = nr_gel_params{};
    int nr = std::sscanf(line.c_str(), "%lf %lf %lf %lf %lf", &g.ks, &g.kb, &g.bmod, &g.smod,
                         &g.dilation);
    if (nr < 3 || g.ks < 0 || g.kb < 0 || g.bmod < 0 || g.smod < 0 || g.dilation < 0) {
        err = "Error: invalid format for mechanical attributes.";
        return NR_E_BATCH;
    }
    return NR_OK;
}

void log_mesh_result(FILE *lf, const nr_report &r, double fcut)
{
    std::printf("Steps: %d, RMS Force: %12.8le\n", r.steps, r.rms_force);
    if (!lf) return;
    std::fprintf(lf, r.rms_force <= fcut ? "Convergence reached: " : "Convergence not reached: ");
    std::fprintf(lf, "Steps: %d, RMS Force: %12.8le\n", r.steps, r.rms_force);
}

// do_relaxation_run :1095-1391 on one composite.  Compression: the first
// step applies the minimum, each later one the increment, the height shrinking
// by each (:1283-1327); shear likewise from the minimum (:1335-1377).
int run_mesh_protocol(const JobFiles &jf, const MeshBatchSpec &s, bool strain_mode,
                      int device, std::string &err)
{
    const BatchSpec &b = s.b;
    if ((b.cmax > 0 && !(b.cinc > 0)) || (b.smax > 0 && !(b.sinc > 0))) {
        err = "a ramp with a positive maximum needs a positive increment";
        return NR_E_BATCH;
    }
    clock_t ct = std::clock();
    nr_mesh *mp = nullptr;
    int rc = nr_mesh_read(jf.dat.c_str(), &mp);
    if (rc) {
        err = nr_last_error();
        return rc;
    }
    std::unique_ptr<nr_mesh, void (*)(nr_mesh *)> mesh(mp, nr_mesh_free);
    nr_composite *cp = nullptr;
    rc = nr_composite_create(mesh.get(), &s.gel, device, &cp);
    if (rc) {
        err = nr_last_error();
        return rc;
    }
    std::unique_ptr<nr_composite, void (*)(nr_composite *)> comp(cp, nr_composite_destroy);
    const int n = mesh->m.net.n;
    double range = comp->setup.range;
    const double smult = strain_mode ? range : 1.0;
    const int digits = dec_digits(1 + (b.sinc > 0 ? (int)((b.smax - b.smin) / b.sinc) : 0));
    nr_relax_params p{};
    p.alpha0 = b.fp[0];
    p.f_alpha = b.fp[1];
    p.dt_max = b.fp[2];
    p.f_inc = b.fp[3];
    p.f_dec = b.fp[4];
    p.nmin = b.nmin;
    p.mass = b.mass;
    p.fcut = b.fcut;
    p.max_steps = b.nmax;
    p.criterion = NR_CRIT_RMS;
    p.method = NR_METHOD_FIRE;
    std::cout << "Preferred dilation: " << comp->setup.pref_d << "\n";
    std::cout.flush();
    FILE *lf = std::fopen(jf.log.c_str(), "w");
    FILE *ef = std::fopen(jf.edata.c_str(), "w");
    std::vector<double> pos(2 * (size_t)n);
    nr_report r{};
    double eg = 0.0, cur = 0.0;
    double mn = b.cmin * smult, mx = b.cmax * smult, inc = b.cinc * smult;
    if (mx > 0) {
        rc = nr_composite_affine(cp, range, 0.0, -mn);
        range -= mn;
        for (cur = mn; rc == NR_OK && cur <= mx + 1e-10; cur += inc) {
            rc = nr_composite_relax(cp, &p, &r, &eg);
            if (rc) break;
            log_mesh_result(lf, r, p.fcut);
            if (ef)
                std::fprintf(ef, "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le\n", 0., cur,
                             r.e_stretch, r.e_bend, eg, r.e_stretch + r.e_bend + eg);
            if (cur < mx) {
                rc = nr_composite_affine(cp, range, 0.0, -inc);
                range -= inc;
            }
        }
    }
    const double max_comp = mx > mn ? cur - inc : mx;
    mn = b.smin * smult;
    mx = b.smax * smult;
    inc = b.sinc * smult;
    int count = 1;
    if (rc == NR_OK && mx > 0) {
        rc = nr_composite_affine(cp, range, mn, 0.0);
        for (cur = mn; rc == NR_OK && cur <= mx + 1e-10; cur += inc) {
            rc = nr_composite_relax(cp, &p, &r, &eg);
            if (rc) break;
            log_mesh_result(lf, r, p.fcut);
            if (ef)
                std::fprintf(ef, "%12.8lf\t%12.8lf\t%12.8le\t%12.8le\t%12.8le\t%12.8le\n", cur,
                             max_comp, r.e_stretch, r.e_bend, eg, r.e_stretch + r.e_bend + eg);
            rc = nr_composite_get_positions(cp, pos.data());
            if (rc) break;
            FILE *pf = std::fopen(report_name(jf.disp, digits, count++, "pos").c_str(), "w");
            if (pf) {
                for (int i = 0; i < n; i++)
                    std::fprintf(pf, "%2.10le\t%2.10le\n", pos[2 * i], pos[2 * i + 1]);
                std::fclose(pf);
            }
            if (cur < mx) rc = nr_composite_affine(cp, range, inc, 0.0);
        }
    }
    if (rc) err = nr_last_error();
    ct = std::clock() - ct;
    if (lf) {
        std::fprintf(lf, "Elapsed time: %f seconds\n", ((float)ct) / CLOCKS_PER_SEC);
        std::fclose(lf);
    }
    if (ef) std::fclose(ef);
    return rc;
}

}  // namespace
}  // namespace nr

// relaxation_batch :1393-1421: every <base>*.nmc of the batch's directory.
extern "C" int nr_mesh_relaxation_batch(const char *batch_file, int strain_mode, int device)
{
    return nr::guarded(NR_E_BATCH, [&]() -> int {
        std::string err;
        nr::MeshBatchSpec s;
        if (!batch_file) return NR_E_ARG;
        int rc = nr::parse_mesh_batch(batch_file, s, err);
        if (rc) {
            std::fprintf(stderr, "%s\nParsing of the batch file failed.\n", err.c_str());
            return rc;
        }
        std::vector<nr::JobFiles> jobs;
        rc = nr::file_lists(s.b.base, s.b.tag, jobs, err, "nmc");
        if (rc) {
            std::fprintf(stderr, "%s\nFile lists could not be read.\n", err.c_str());
            return rc;
        }
        int first_err = NR_OK;
        for (const nr::JobFiles &j : jobs) {
            rc = nr::run_mesh_protocol(j, s, strain_mode != 0, device, err);
            if (rc) {
                std::fprintf(stderr, "%s: %s\n", j.dat.c_str(), err.c_str());
                if (!first_err) first_err = rc;
            }
        }
        return first_err;
    });
}
