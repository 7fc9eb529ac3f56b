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
#include <cmath>
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
    std::sort(T.begin(), T.end(), [](const Trip &x, const Trip &y) {
        return x.r != y.r ? x.r < y.r : x.c < y.c;
    });
    Csr K;
    K.n = 2 * N.n;
    K.rowptr.assign(K.n + 1, 0);
    for (size_t q = 0; q < T.size(); q++) {
        if (q > 0 && T[q].r == T[q - 1].r && T[q].c == T[q - 1].c) {
            K.val.back() += T[q].v;
        } else {
            K.col.push_back(T[q].c);
            K.val.push_back(T[q].v);
            K.rowptr[T[q].r + 1]++;
        }
    }
    for (int r = 0; r < K.n; r++) K.rowptr[r + 1] += K.rowptr[r];
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

}  // namespace

int nr_linear_response(const nr_network *net, double ks, double kb, double mu, double shear,
                       double compression, int affine, double tol, int max_iter, double *u,
                       nr_linear_report *rep)
{
    return nr::guarded(NR_E_ARG, [&]() -> int {
        if (!net || !rep || !(tol >= 0.0) || max_iter < 0) return fail(NR_E_ARG, "invalid arguments");
        const nr::Network &N = net->net;
        const bool le = N.height > 0;
        if (le && !affine)
            return fail(NR_E_ARG, "a periodic box has no top row to displace: use the affine field");
        if ((kb > 0 || !N.tkb.empty()) && !N.triples_built)
            return fail(NR_E_ARG, "bending needs the triples (nr_network_build_triples)");
        const int n2 = 2 * N.n;
        const Csr K = assemble(N, ks, kb, mu);

        // imposed field and its load
        std::vector<double> bc(n2, 0.0), g, kbc;
        std::vector<unsigned char> fr(n2);
        for (int i = 0; i < N.n; i++) {
            const double y = N.pos[2 * i + 1];
            if (affine) {
                bc[2 * i] = shear * (y - N.miny);
                bc[2 * i + 1] = compression * (N.miny - y);
            } else if (y > N.maxy - 1e-10) {
                bc[2 * i] = shear;
                bc[2 * i + 1] = -compression;
            }
            fr[2 * i] = (N.fcode[i] & 1) != 0;
            fr[2 * i + 1] = (N.fcode[i] & 2) != 0;
        }
        spmv(K, bc, g);
        kbc = g;
        const double ix = affine ? shear * N.height : 0.0, iy = affine ? -compression * N.height : 0.0;
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
        std::vector<double> b0(n2);
        for (int k = 0; k < n2; k++) b0[k] = fr[k] ? -g[k] : 0.0;
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
        const double rr0 = dot(b0, b0);
        const double thr2 = (tol * scale) * (tol * scale);
        const bool done0 = !(rr0 > thr2) || max_iter == 0;

        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int rc = NR_OK;
        double sc[nr::kLinScalars] = {rr0, 0.0, thr2, done0 ? 1.0 : 0.0, 0.0, 0.0, 0.0, rr0, 0.0};
        float ms = 0.0f;
        auto cleanup = [&]() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (st) (void)hipStreamDestroy(st);
        };
        // Every copy goes through this solve's own non-blocking stream: other
        // host threads may be capturing their solves' graphs meanwhile, and a
        // copy on the legacy stream would be refused then.
        DevBuf<int> d_rowptr, d_col;
        DevBuf<double> d_val, d_w, d_r, d_p, d_p2, d_q, d_part, d_sc;
        DevBuf<unsigned char> d_free;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        auto put = [&](auto &buf, const auto &host) {
            if (e != hipSuccess) return;
            e = buf.alloc(host.size());
            if (e == hipSuccess && !host.empty())
                e = hipMemcpyAsync(buf.p, host.data(), sizeof(host[0]) * host.size(),
                                   hipMemcpyHostToDevice, st);
        };
        auto zeros = [&](DevBuf<double> &buf, size_t n) {
            if (e != hipSuccess) return;
            e = buf.alloc(n);
            if (e == hipSuccess) e = hipMemsetAsync(buf.p, 0, sizeof(double) * n, st);
        };
        const std::vector<double> sc0(sc, sc + nr::kLinScalars);
        put(d_rowptr, K.rowptr);
        put(d_col, K.col);
        put(d_val, K.val);
        put(d_free, fr);
        put(d_r, b0);
        put(d_sc, sc0);
        zeros(d_w, n2);
        zeros(d_p, n2);  // the first direction is r (beta = 0)
        zeros(d_p2, n2);
        zeros(d_q, n2);
        zeros(d_part, 2 * nr::kLinBlocks);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        // a chunk of iterations as one graph (kernels only); replays stop doing
        // work once the device raises its done flag
        constexpr int kChunk = 32;  // even: the parity of iteration k repeats per chunk
        double *const pbuf[2] = {d_p.p, d_p2.p};
        if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (int c = 0; c < kChunk && e == hipSuccess; c++)
                e = (hipError_t)nr::launch_cg_iteration(n2, d_rowptr.p, d_col.p, d_val.p, d_free.p,
                                                        d_w.p, d_r.p, pbuf, d_q.p, d_part.p, d_sc.p,
                                                        c, max_iter, st);
            if (e == hipSuccess) e = (hipError_t)nr::launch_cg_chunk_end(d_sc.p, kChunk, st);
            hipError_t e2 = hipStreamEndCapture(st, &graph);
            if (e == hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        while (e == hipSuccess && sc[3] == 0.0 && sc[8] == 0.0) {
            e = hipGraphLaunch(exec, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(sc, d_sc.p, sizeof(sc), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
        }
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        std::vector<double> w(n2);
        if (e == hipSuccess)
            e = hipMemcpyAsync(w.data(), d_w.p, sizeof(double) * n2, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        cleanup();
        if (e != hipSuccess) return nr::hip_fail(e, "linear solve");

        // u, the true residual of the reduced system, the linearised energies
        std::vector<double> kw;
        spmv(K, w, kw);
        double res = 0.0;
        for (int k = 0; k < n2; k++)
            if (fr[k]) res += (kw[k] + g[k]) * (kw[k] + g[k]);
        std::vector<double> uf(n2);
        for (int k = 0; k < n2; k++) uf[k] = bc[k] + w[k];
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
        rep->status = !(sc[7] > thr2) ? NR_CONVERGED : NR_NOT_CONVERGED;
        rep->device_ms = ms;
        if (u) std::copy(uf.begin(), uf.end(), u);
        return rc;
    });
}
