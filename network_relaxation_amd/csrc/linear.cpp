// Small-strain linear response of one network (the reference's
// sparse_relax.cpp), solved on the device.
//
// Host: the stiffness K is assembled from the reference's element blocks
// (calc_stretching_elements :286-322, ks d d^T / l^3; calc_bending_elements
// :325-386, the Hessian of the linearised bending energy; calc_shearing_elements
// :390-428, the gel term) into CSR over all 2N DOFs; the load g = K bc is the
// force of the imposed field bc (affine_displacement :527-537 or
// add_displacement :513-523), in a Lees-Edwards box plus the elements that
// reach a node's image across the y boundary.  Device: conjugate gradients on
// the reduced system K_rr w = -g_r (get_lhs_rhs :560-605; the reference
// factorises it by sparse QR) from w = 0, so that w is the minimum-norm
// solution; the iterations run as a replayed HIP graph with device-side
// stopping.  Host again: u = bc + w, the residual 2-norm |K_rr w + g_r| and
// the linearised energies (get_stretch_shear_energy :49-70,
// get_bending_energy :73-96), as do_relaxation_run reports them (:676-881).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace {

using nr::DevBuf;
using nr::fail;

struct Trip {
    int r, c;
    double v;
};

struct Csr {
    int n = 0;
    std::vector<int> rowptr, col;
    std::vector<double> val;
};

void bond_geom(const nr::Network &N, int b, double &dx, double &dy)
{
    const int i = N.bonds[2 * b], j = N.bonds[2 * b + 1];
    dx = N.pos[2 * j] + N.width * N.bwrap[2 * b] - N.pos[2 * i];
    dy = N.pos[2 * j + 1] + N.height * N.bwrap[2 * b + 1] - N.pos[2 * i + 1];
}

void arm_geom(const nr::Network &N, int t, double &dx, double &dy)
{
    const int h = N.tri[3 * t], a = N.tri[3 * t + 1];
    dx = N.pos[2 * a] + N.width * N.twrap[4 * t] - N.pos[2 * h];
    dy = N.pos[2 * a + 1] + N.height * N.twrap[4 * t + 1] - N.pos[2 * h + 1];
}

double tri_kappa(const nr::Network &N, double kb, int t)
{
    return N.tkb.empty() ? kb : N.tkb[t];
}

void put(std::vector<Trip> &T, int r, int c, double v)
{
    T.push_back({r, c, v});
    if (r != c) T.push_back({c, r, v});
}

Csr assemble(const nr::Network &N, double ks, double kb, double mu)
{
    std::vector<Trip> T;
    T.reserve((size_t)N.e() * 24 + (size_t)N.t() * 36);
    for (int b = 0; ks > 0 && b < N.e(); b++) {
        const int xi = 2 * N.bonds[2 * b], xj = 2 * N.bonds[2 * b + 1];
        double dx, dy;
        bond_geom(N, b, dx, dy);
        const double l = std::sqrt(dx * dx + dy * dy);
        const double s = ks / (l * l * l);
        const double xx = s * dx * dx, xy = s * dx * dy, yy = s * dy * dy;
        put(T, xi, xi, xx);
        put(T, xj, xj, xx);
        put(T, xi, xj, -xx);
        put(T, xi + 1, xi + 1, yy);
        put(T, xj + 1, xj + 1, yy);
        put(T, xi + 1, xj + 1, -yy);
        put(T, xi, xi + 1, xy);
        put(T, xj, xj + 1, xy);
        put(T, xi, xj + 1, -xy);
        put(T, xi + 1, xj, -xy);
    }
    // bending: kb/|r|^2 g g^T, g = d(ddu x r)/du with ddu = u_a + u_b - 2 u_h
    static const double gx[6] = {-2.0, 0.0, 1.0, 0.0, 1.0, 0.0};  // times dy
    static const double gy[6] = {0.0, 2.0, 0.0, -1.0, 0.0, -1.0}; // times dx
    for (int t = 0; t < N.t(); t++) {
        const double k = tri_kappa(N, kb, t);
        if (!(k > 0)) continue;
        const int h = 2 * N.tri[3 * t], a = 2 * N.tri[3 * t + 1], c = 2 * N.tri[3 * t + 2];
        double dx, dy;
        arm_geom(N, t, dx, dy);
        const double l2 = dx * dx + dy * dy;
        const double xx = k * dx * dx / l2, yy = k * dy * dy / l2, xy = k * dx * dy / l2;
        const int dof[6] = {h, h + 1, a, a + 1, c, c + 1};
        for (int p = 0; p < 6; p++)
            for (int q = p; q < 6; q++) {
                const double v = gx[p] * gx[q] * yy + (gx[p] * gy[q] + gy[p] * gx[q]) * xy +
                                 gy[p] * gy[q] * xx;
                if (v != 0.0) put(T, dof[p], dof[q], v);
            }
    }
    for (int b = 0; mu > 0 && b < N.e(); b++) {
        const int xi = 2 * N.bonds[2 * b], xj = 2 * N.bonds[2 * b + 1];
        double dx, dy;
        bond_geom(N, b, dx, dy);
        const double inv = 1.0 / N.rest[b] / N.rest[b];
        const double mx = mu * (1.0 - dx * dx * inv), my = mu * (1.0 - dy * dy * inv);
        put(T, xi, xi, mx);
        put(T, xj, xj, mx);
        put(T, xi, xj, -mx);
        put(T, xi + 1, xi + 1, my);
        put(T, xj + 1, xj + 1, my);
        put(T, xi + 1, xj + 1, -my);
    }
    // bucket the entries by row (in insertion order), then per row order them
    // by column (stable) and sum equal positions in insertion order
    Csr K;
    K.n = 2 * N.n;
    std::vector<int> start(K.n + 1, 0);
    for (const Trip &t : T) start[t.r + 1]++;
    for (int r = 0; r < K.n; r++) start[r + 1] += start[r];
    std::vector<std::pair<int, double>> ent(T.size());
    {
        std::vector<int> at(start.begin(), start.end() - 1);
        for (const Trip &t : T) ent[at[t.r]++] = {t.c, t.v};
    }
    K.rowptr.assign(K.n + 1, 0);
    K.col.reserve(T.size() / 2);
    K.val.reserve(T.size() / 2);
    for (int r = 0; r < K.n; r++) {
        auto lo = ent.begin() + start[r], hi = ent.begin() + start[r + 1];
        std::stable_sort(lo, hi, [](const std::pair<int, double> &x,
                                    const std::pair<int, double> &y) { return x.first < y.first; });
        for (auto q = lo; q != hi; ++q) {
            if (q != lo && q->first == (q - 1)->first) {
                K.val.back() += q->second;
            } else {
                K.col.push_back(q->first);
                K.val.push_back(q->second);
            }
        }
        K.rowptr[r + 1] = (int)K.col.size();
    }
    return K;
}

void spmv(const Csr &K, const std::vector<double> &x, std::vector<double> &y)
{
    y.assign(K.n, 0.0);
    for (int r = 0; r < K.n; r++) {
        double s = 0.0;
        for (int q = K.rowptr[r]; q < K.rowptr[r + 1]; q++) s = std::fma(K.val[q], x[K.col[q]], s);
        y[r] = s;
    }
}

double dot(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size(); i++) s = std::fma(a[i], b[i], s);
    return s;
}

// Runs f(0) .. f(count - 1) on up to 16 host threads (each index once); the
// first exception any call throws is rethrown here once all have finished.
template <class F>
void parallel_for(int count, F f)
{
    const int nt = std::max(1, std::min({count, 16, (int)std::thread::hardware_concurrency()}));
    std::atomic<int> next{0};
    std::exception_ptr first;
    std::mutex mu;
    auto work = [&]() {
        for (int i; (i = next.fetch_add(1)) < count;) {
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                if (!first) first = std::current_exception();
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
    if (first) std::rethrow_exception(first);
}

// Host side of one system: the imposed field, its load and the reduced
// right-hand side, and afterwards the report.
struct Prep {
    const nr_linear_job *job = nullptr;
    const Csr *K = nullptr;
    std::vector<double> bc, g, b0;
    std::vector<unsigned char> fr;
    double ix = 0.0, iy = 0.0, rr0 = 0.0, thr2 = 0.0;
    bool done0 = false;
};

int check_job(const nr_linear_job &jb)
{
    if (!jb.net || !(jb.tol >= 0.0) || jb.max_iter < 0 || !(jb.ks >= 0.0) || !(jb.kb >= 0.0) ||
        !(jb.mu >= 0.0))
        return fail(NR_E_ARG, "invalid arguments");
    const nr::Network &N = jb.net->net;
    if (N.height > 0 && !jb.affine)
        return fail(NR_E_ARG, "a periodic box has no top row to displace: use the affine field");
    if ((jb.kb > 0 || !N.tkb.empty()) && !N.triples_built)
        return fail(NR_E_ARG, "bending needs the triples (nr_network_build_triples)");
    return NR_OK;
}

void prepare(const nr_linear_job &jb, const Csr &K, Prep &P)
{
    const nr::Network &N = jb.net->net;
    const bool le = N.height > 0;
    const int n2 = 2 * N.n;
    const double ks = jb.ks, kb = jb.kb, mu = jb.mu, shear = jb.shear, compression = jb.compression;
    const bool affine = jb.affine != 0;
    P.job = &jb;
    P.K = &K;
    // imposed field and its load
    std::vector<double> &bc = P.bc, &g = P.g;
    bc.assign(n2, 0.0);
    P.fr.resize(n2);
    for (int i = 0; i < N.n; i++) {
        const double y = N.pos[2 * i + 1];
        if (affine) {
            bc[2 * i] = shear * (y - N.miny);
            bc[2 * i + 1] = compression * (N.miny - y);
        } else if (y > N.maxy - 1e-10) {
            bc[2 * i] = shear;
            bc[2 * i + 1] = -compression;
        }
        P.fr[2 * i] = (N.fcode[i] & 1) != 0;
        P.fr[2 * i + 1] = (N.fcode[i] & 2) != 0;
    }
    spmv(K, bc, g);
    const std::vector<double> kbc = g;
    const double ix = affine ? shear * N.height : 0.0, iy = affine ? -compression * N.height : 0.0;
    P.ix = ix;
    P.iy = iy;
    if (le && affine) {
        for (int b = 0; b < N.e(); b++) {
            const int wy = N.bwrap[2 * b + 1];
            if (!wy) continue;
            const int i = N.bonds[2 * b], j = N.bonds[2 * b + 1];
            double dx, dy;
            bond_geom(N, b, dx, dy);
            const double l = std::sqrt(dx * dx + dy * dy);
            const double sk = ks / (l * l * l) * (dx * ix * wy + dy * iy * wy);
            g[2 * j] += sk * dx;
            g[2 * j + 1] += sk * dy;
            g[2 * i] -= sk * dx;
            g[2 * i + 1] -= sk * dy;
            if (mu > 0) {
                const double inv = 1.0 / N.rest[b] / N.rest[b];
                const double mx = mu * (1.0 - dx * dx * inv) * ix * wy;
                const double my = mu * (1.0 - dy * dy * inv) * iy * wy;
                g[2 * j] += mx;
                g[2 * j + 1] += my;
                g[2 * i] -= mx;
                g[2 * i + 1] -= my;
            }
        }
        for (int t = 0; t < N.t(); t++) {
            const int wya = N.twrap[4 * t + 1], wyb = N.twrap[4 * t + 3];
            if (!wya && !wyb) continue;
            const double k = tri_kappa(N, kb, t);
            const int h = N.tri[3 * t], a = N.tri[3 * t + 1], c = N.tri[3 * t + 2];
            double dx, dy;
            arm_geom(N, t, dx, dy);
            const double l2 = dx * dx + dy * dy;
            const double f = k * ((ix * (wya + wyb)) * dy - (iy * (wya + wyb)) * dx) / l2;
            g[2 * h] += -2.0 * dy * f;
            g[2 * h + 1] += 2.0 * dx * f;
            g[2 * a] += dy * f;
            g[2 * a + 1] += -dx * f;
            g[2 * c] += dy * f;
            g[2 * c + 1] += -dx * f;
        }
    }
    // right-hand side; in a periodic box every node is free and the two
    // translations are null vectors of K, which the load is orthogonal to up
    // to round-off
    std::vector<double> &b0 = P.b0;
    b0.resize(n2);
    for (int k = 0; k < n2; k++) b0[k] = P.fr[k] ? -g[k] : 0.0;
    if (le && N.n > 0) {
        double mx = 0.0, my = 0.0;
        for (int i = 0; i < N.n; i++) {
            mx += b0[2 * i];
            my += b0[2 * i + 1];
        }
        mx /= N.n;
        my /= N.n;
        for (int i = 0; i < N.n; i++) {
            b0[2 * i] -= mx;
            b0[2 * i + 1] -= my;
        }
    }
    const double scale = std::max(std::sqrt(dot(g, g)), std::sqrt(dot(kbc, kbc)));
    P.rr0 = dot(b0, b0);
    P.thr2 = (jb.tol * scale) * (jb.tol * scale);
    P.done0 = !(P.rr0 > P.thr2) || jb.max_iter == 0;
}

// u, the true residual of the reduced system, the linearised energies
void finish(const Prep &P, const std::vector<double> &w, const double *sc, float ms,
            nr_linear_report *rep)
{
    const nr_linear_job &jb = *P.job;
    const nr::Network &N = jb.net->net;
    const int n2 = 2 * N.n;
    const double ks = jb.ks, kb = jb.kb, mu = jb.mu, ix = P.ix, iy = P.iy;
    std::vector<double> kw;
    spmv(*P.K, w, kw);
    double res = 0.0;
    for (int k = 0; k < n2; k++)
        if (P.fr[k]) res += (kw[k] + P.g[k]) * (kw[k] + P.g[k]);
    std::vector<double> uf(n2);
    for (int k = 0; k < n2; k++) uf[k] = P.bc[k] + w[k];
    double es = 0.0, eb = 0.0;
    for (int b = 0; b < N.e(); b++) {
        const int i = N.bonds[2 * b], j = N.bonds[2 * b + 1];
        const int wy = N.bwrap[2 * b + 1];
        double dx, dy;
        bond_geom(N, b, dx, dy);
        const double dux = uf[2 * j] + ix * wy - uf[2 * i];
        const double duy = uf[2 * j + 1] + iy * wy - uf[2 * i + 1];
        const double de = (dux * dx + duy * dy) / N.rest[b];
        es += 0.5 * ks * de * de;
        es += 0.5 * mu * (dux * dux + duy * duy - de * de);
    }
    for (int t = 0; t < N.t(); t++) {
        const int h = N.tri[3 * t], a = N.tri[3 * t + 1], c = N.tri[3 * t + 2];
        const int wya = N.twrap[4 * t + 1], wyb = N.twrap[4 * t + 3];
        double dx, dy;
        arm_geom(N, t, dx, dy);
        const double l2 = dx * dx + dy * dy;
        const double dux = uf[2 * a] + ix * wya + uf[2 * c] + ix * wyb - 2 * uf[2 * h];
        const double duy = uf[2 * a + 1] + iy * wya + uf[2 * c + 1] + iy * wyb - 2 * uf[2 * h + 1];
        const double cr = dux * dy - duy * dx;
        eb += 0.5 * tri_kappa(N, kb, t) * cr * cr / l2;
    }
    rep->e_stretch = es;
    rep->e_bend = eb;
    rep->residual = std::sqrt(res);
    rep->iterations = (int)sc[6];
    rep->status = !(sc[7] > P.thr2) ? NR_CONVERGED : NR_NOT_CONVERGED;
    rep->device_ms = ms;
    if (jb.u) std::copy(uf.begin(), uf.end(), jb.u);
}

// Streams a batch's systems are spread over, each group of systems running
// its chunks of iterations as its own graph on its own stream.  One: four
// sampled networks (L 22-162) took 343 ms of device time on four streams
// and 273 ms as one group (each iteration's two launches serve them all).
constexpr int kLinStreams = 1;

// One group of a batch's systems on the device: the matrices its systems
// use, their vectors, segments and scalars, its stream and chunk graph.
struct DevGroup {
    std::vector<int> jobs;      // the batch's job indices
    std::vector<nr::LinSeg> segs;
    std::vector<double> sc;     // host copy of the scalars
    double *h_sc = nullptr;     // pinned staging of the scalars' copy back
    long long rows = 0;
    int blocks = 0;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf<int> d_rowptr, d_col, d_seg_of;
    DevBuf<nr::LinSeg> d_segs;
    DevBuf<double> d_val, d_w, d_r, d_p, d_p2, d_q, d_part, d_sc;
    DevBuf<unsigned char> d_free;
    float ms = 0.0f;
    std::vector<double> w;

    ~DevGroup()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        if (h_sc) (void)hipHostFree(h_sc);
    }

    bool running() const
    {
        for (size_t i = 0; i < jobs.size(); i++)
            if (sc[nr::kLinScalars * i + 3] == 0.0 && sc[nr::kLinScalars * i + 8] == 0.0) return true;
        return false;
    }

    // Builds and uploads the group's arrays and captures its chunk graph.
    // Every copy goes through the group's own non-blocking stream: other host
    // threads may be capturing their solves' graphs meanwhile, and a copy on
    // the legacy stream would be refused then.
    hipError_t setup(const std::vector<Prep> &preps, const std::vector<Csr> &mats,
                     const std::vector<int> &mat_of, const nr_linear_job *all)
    {
        std::vector<int> rowptr, col, seg_of, mat_rp(mats.size(), -1), mat_fr(mats.size(), -1);
        std::vector<long long> mat_nz(mats.size(), -1);
        std::vector<double> val, r0;
        std::vector<unsigned char> fr;
        segs.resize(jobs.size());
        for (size_t k = 0; k < jobs.size(); k++) {
            const int i = jobs[k], m = mat_of[i];
            const Prep &P = preps[i];
            if (mat_rp[m] < 0) {
                mat_rp[m] = (int)rowptr.size();
                mat_nz[m] = (long long)col.size();
                rowptr.insert(rowptr.end(), mats[m].rowptr.begin(), mats[m].rowptr.end());
                col.insert(col.end(), mats[m].col.begin(), mats[m].col.end());
                val.insert(val.end(), mats[m].val.begin(), mats[m].val.end());
            }
            nr::LinSeg &S = segs[k];
            const int n = (int)P.b0.size();
            if (rows + n > 0x7fffffffll) return hipErrorInvalidValue;
            S.row0 = (int)rows;
            S.n = n;
            S.blk0 = (int)seg_of.size();
            S.nblk = nr::lin_segment_blocks(n);
            S.rp0 = mat_rp[m];
            S.fr0 = (int)fr.size();
            S.max_iter = all[i].max_iter;
            S.pad = 0;
            S.nz0 = mat_nz[m];
            seg_of.insert(seg_of.end(), S.nblk, (int)k);
            fr.insert(fr.end(), P.fr.begin(), P.fr.end());
            r0.insert(r0.end(), P.b0.begin(), P.b0.end());
            rows += n;
            const double s0[nr::kLinScalars] = {P.rr0, 0.0, P.thr2, P.done0 ? 1.0 : 0.0, 0.0,
                                                0.0,   0.0, P.rr0,  0.0};
            sc.insert(sc.end(), s0, s0 + nr::kLinScalars);
        }
        blocks = (int)seg_of.size();
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        // pinned, so that a group's copy back does not hold the host thread
        // (and with it the other groups' launches) until its chunk is done
        if (e == hipSuccess)
            e = hipHostMalloc((void **)&h_sc, sizeof(double) * sc.size(), hipHostMallocDefault);
        auto put = [&](auto &buf, const auto &host) {
            if (e != hipSuccess) return;
            e = buf.alloc(std::max<size_t>(host.size(), 1));
            if (e == hipSuccess && !host.empty())
                e = hipMemcpyAsync(buf.p, host.data(), sizeof(host[0]) * host.size(),
                                   hipMemcpyHostToDevice, st);
        };
        auto zeros = [&](DevBuf<double> &buf, size_t n) {
            if (e != hipSuccess) return;
            n = std::max<size_t>(n, 1);
            e = buf.alloc(n);
            if (e == hipSuccess) e = hipMemsetAsync(buf.p, 0, sizeof(double) * n, st);
        };
        put(d_rowptr, rowptr);
        put(d_col, col);
        put(d_val, val);
        put(d_free, fr);
        put(d_segs, segs);
        put(d_seg_of, seg_of);
        put(d_r, r0);
        put(d_sc, sc);
        zeros(d_w, (size_t)rows);
        zeros(d_p, (size_t)rows);  // the first direction is r (beta = 0)
        zeros(d_p2, (size_t)rows);
        zeros(d_q, (size_t)rows);
        zeros(d_part, 2 * (size_t)blocks);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        // a chunk of iterations as one graph (kernels only); replays stop
        // doing work for a system once the device raises its done flag
        double *const pbuf[2] = {d_p.p, d_p2.p};
        if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (int c = 0; c < kChunk && e == hipSuccess; c++)
                e = (hipError_t)nr::launch_cg_iteration(d_segs.p, d_seg_of.p, blocks, d_rowptr.p,
                                                        d_col.p, d_val.p, d_free.p, d_w.p, d_r.p,
                                                        pbuf, d_q.p, d_part.p, d_sc.p, c, st);
            if (e == hipSuccess)
                e = (hipError_t)nr::launch_cg_chunk_end(d_sc.p, (int)jobs.size(), kChunk, st);
            hipError_t e2 = hipStreamEndCapture(st, &graph);
            if (e == hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        return e;
    }

    // One chunk of iterations and the scalars' copy back (asynchronous; the
    // host copy is read after the stream's synchronisation, in sync()).
    hipError_t chunk()
    {
        hipError_t e = hipGraphLaunch(exec, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h_sc, d_sc.p, sizeof(double) * sc.size(), hipMemcpyDeviceToHost, st);
        return e;
    }

    hipError_t sync()
    {
        const hipError_t e = hipStreamSynchronize(st);
        if (e == hipSuccess) std::copy(h_sc, h_sc + sc.size(), sc.begin());
        return e;
    }

    hipError_t collect()
    {
        hipError_t e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        w.assign((size_t)rows, 0.0);
        if (e == hipSuccess && rows > 0)
            e = hipMemcpyAsync(w.data(), d_w.p, sizeof(double) * rows, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        return e;
    }

    static constexpr int kChunk = 32;  // even: the parity of iteration k repeats per chunk
};

// The systems of a batch on the device: their matrices (one per network and
// stiffness, built with the loads on several host threads), then CG on every
// system at once, in up to kLinStreams groups balanced by rows, each group's
// chunks of iterations replayed as its own graph until all its systems are
// done.
int solve_batch(int count, const nr_linear_job *jobs, nr_linear_report *reps)
{
    for (int i = 0; i < count; i++) {
        const int rc = check_job(jobs[i]);
        if (rc) return rc;
    }
    // one matrix per (network, ks, kb, mu)
    std::vector<int> mat_of(count), mat_job;
    for (int i = 0; i < count; i++) {
        const nr_linear_job &a = jobs[i];
        int m = -1;
        for (int j = 0; j < i && m < 0; j++) {
            const nr_linear_job &b = jobs[j];
            if (b.net == a.net && b.ks == a.ks && b.kb == a.kb && b.mu == a.mu) m = mat_of[j];
        }
        if (m < 0) {
            m = (int)mat_job.size();
            mat_job.push_back(i);
        }
        mat_of[i] = m;
    }
    std::vector<Csr> mats(mat_job.size());
    parallel_for((int)mats.size(), [&](int m) {
        const nr_linear_job &a = jobs[mat_job[m]];
        mats[m] = assemble(a.net->net, a.ks, a.kb, a.mu);
    });
    std::vector<Prep> preps(count);
    parallel_for(count, [&](int i) { prepare(jobs[i], mats[mat_of[i]], preps[i]); });

    // groups: the largest systems first, each to the group with the fewest rows
    const int ng = std::min(count, kLinStreams);
    std::vector<DevGroup> groups(ng);
    std::vector<long long> load(ng, 0);
    std::vector<int> order(count);
    for (int i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return preps[a].b0.size() > preps[b].b0.size(); });
    for (int i : order) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        groups[g].jobs.push_back(i);
        load[g] += (long long)preps[i].b0.size() + 1;
    }
    for (DevGroup &G : groups) std::sort(G.jobs.begin(), G.jobs.end());
    hipError_t e = hipSuccess;
    for (DevGroup &G : groups)
        if (e == hipSuccess) e = G.setup(preps, mats, mat_of, jobs);
    // chunks on every group still running, then every such group's scalars
    std::vector<char> live(ng, 1);
    while (e == hipSuccess) {
        bool any = false;
        for (int g = 0; g < ng; g++) {
            live[g] = groups[g].running();
            if (live[g] && e == hipSuccess) {
                e = groups[g].chunk();
                any = true;
            }
        }
        if (!any) break;
        for (int g = 0; g < ng; g++)
            if (live[g] && e == hipSuccess) e = groups[g].sync();
    }
    for (DevGroup &G : groups)
        if (e == hipSuccess) e = G.collect();
    if (e != hipSuccess) return nr::hip_fail(e, "linear solve");
    float ms = 0.0f;
    for (const DevGroup &G : groups) ms = std::max(ms, G.ms);
    for (const DevGroup &G : groups)
        parallel_for((int)G.jobs.size(), [&](int k) {
            const nr::LinSeg &S = G.segs[k];
            const std::vector<double> wi(G.w.begin() + S.row0, G.w.begin() + S.row0 + S.n);
            finish(preps[G.jobs[k]], wi, G.sc.data() + (size_t)nr::kLinScalars * k, ms,
                   reps + G.jobs[k]);
        });
    return NR_OK;
}

}  // namespace

int nr_linear_response(const nr_network *net, double ks, double kb, double mu, double shear,
                       double compression, int affine, double tol, int max_iter, double *u,
                       nr_linear_report *rep)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!rep) return fail(NR_E_ARG, "invalid arguments");
        nr_linear_job jb;
        jb.net = net;
        jb.ks = ks;
        jb.kb = kb;
        jb.mu = mu;
        jb.shear = shear;
        jb.compression = compression;
        jb.affine = affine;
        jb.tol = tol;
        jb.max_iter = max_iter;
        jb.u = u;
        return solve_batch(1, &jb, rep);
    });
}

int nr_linear_response_batch(int count, const nr_linear_job *jobs, nr_linear_report *reps)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (count < 0 || (count > 0 && (!jobs || !reps))) return fail(NR_E_ARG, "invalid arguments");
        if (count == 0) return NR_OK;
        return solve_batch(count, jobs, reps);
    });
}
