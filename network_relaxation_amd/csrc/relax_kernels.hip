// gfx950 kernels of the relax step: one persistent launch relaxes every
// network of a batch.  Workgroups form teams; a team takes networks from a
// queue (largest first) and runs FIRE on each to convergence, each member
// owning a fixed set of spatial tiles.  Per FIRE step and tile:
//   stage the tile's bond/triple entries into LDS (coalesced copy),
//   update own + halo positions into LDS (the deferred FIRE mix, drift and
//   half kick of the previous step, or the affine / Lees-Edwards strain step
//   on the first pass of a relax),
//   gather each own node's force from its entries (no atomics; entries are in
//   element order so the sum order is fixed), half kick, partial reductions.
// One team barrier per step combines the partials (P = v.F, |v|^2, |F|^2,
// max |F_node|^2) in member order, so every member takes the same FIRE
// decision.  References: run_fire_md large_strain_relax.cu:596-762,
// axial_forces :386-414, bending_forces :417-469, energies :472-523,
// affine_displace :527-537.
#include <hip/hip_runtime.h>

#include "device.hpp"

#pragma clang fp contract(off)

namespace nr {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

enum PassKind { kPassInit = 0, kPassStep = 1, kPassEnergy = 2 };
enum MixKind { kMixNone = 0, kMixFire = 1, kMixZero = 2 };

struct Red4 {
    double a, b, c, m;  // sums a, b, c and max m
};

struct StepScal {
    double dt, dtsq, dtm;   // drift / kick of this pass
    double alpha_old, cmix; // deferred FIRE mix of the previous step
    int mix;
    double pend_x, pend_y;  // affine step (init pass)
};

struct NetCtx {
    int node_base, n_nodes, tile_base, n_tiles;
    double W, H, shift, miny, yrange, ks, kb;
};

__device__ inline int wrap_of(uint32_t v) { return (int)(v & 0xF) - kWrapBias; }

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Bounded team barrier, executed by thread 0 of each member workgroup after
// a __syncthreads().  Returns false when the wait bound expired.
__device__ bool team_arrive(unsigned *cnt, unsigned *genp, unsigned g,
                            unsigned &gen, long long timeout)
{
    unsigned target = gen + 1;
    unsigned old = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == g - 1) {
        __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(genp, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        long long t0 = wall_clock64();
        while (__hip_atomic_load(genp, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != target) {
            __builtin_amdgcn_s_sleep(2);
            if (wall_clock64() - t0 > timeout) return false;
        }
    }
    gen = target;
    return true;
}

struct Team {
    int id, member, g;
    unsigned gen;
};

// Whole-workgroup team barrier; all threads return the same verdict.
__device__ bool team_sync(const KArgs &A, Team &T, int *s_flag)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        bool ok = team_arrive(A.bar + T.id * 64, A.bar + T.id * 64 + 32, (unsigned)T.g, T.gen,
                              A.timeout_ticks);
        if (!ok) __hip_atomic_store(A.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_flag = ok ? 1 : 0;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return *s_flag != 0;
}

// Block + team reduction of a Red4 in a fixed order (wave butterfly, waves
// in order, members in order).  Every thread gets the result.
__device__ bool reduce4(const KArgs &A, Team &T, Red4 r, Red4 &out, double *s_red,
                        int *s_flag)
{
    r.a = wave_sum(r.a);
    r.b = wave_sum(r.b);
    r.c = wave_sum(r.c);
    r.m = wave_max(r.m);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[4 * wave + 0] = r.a;
        s_red[4 * wave + 1] = r.b;
        s_red[4 * wave + 2] = r.c;
        s_red[4 * wave + 3] = r.m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Red4 t = {0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < kWaves; w++) {
            t.a += s_red[4 * w + 0];
            t.b += s_red[4 * w + 1];
            t.c += s_red[4 * w + 2];
            t.m = fmax(t.m, s_red[4 * w + 3]);
        }
        s_red[4 * kWaves + 0] = t.a;
        s_red[4 * kWaves + 1] = t.b;
        s_red[4 * kWaves + 2] = t.c;
        s_red[4 * kWaves + 3] = t.m;
        if (T.g > 1) {
            double *slot = A.partials + ((size_t)T.id * T.g + T.member) * 4;
            slot[0] = t.a;
            slot[1] = t.b;
            slot[2] = t.c;
            slot[3] = t.m;
        }
    }
    if (T.g > 1) {
        if (!team_sync(A, T, s_flag)) return false;
        if (threadIdx.x == 0) {
            Red4 t = {0.0, 0.0, 0.0, 0.0};
            const double *base = A.partials + (size_t)T.id * T.g * 4;
            for (int m = 0; m < T.g; m++) {
                t.a += __builtin_nontemporal_load(base + 4 * m + 0);
                t.b += __builtin_nontemporal_load(base + 4 * m + 1);
                t.c += __builtin_nontemporal_load(base + 4 * m + 2);
                t.m = fmax(t.m, __builtin_nontemporal_load(base + 4 * m + 3));
            }
            s_red[4 * kWaves + 0] = t.a;
            s_red[4 * kWaves + 1] = t.b;
            s_red[4 * kWaves + 2] = t.c;
            s_red[4 * kWaves + 3] = t.m;
        }
    }
    __syncthreads();
    out.a = s_red[4 * kWaves + 0];
    out.b = s_red[4 * kWaves + 1];
    out.c = s_red[4 * kWaves + 2];
    out.m = s_red[4 * kWaves + 3];
    __syncthreads();
    return true;
}

struct Lds {
    double2 *pos;
    double *blen;
    uint64_t *tcode;
    double *tkb;
    uint32_t *bcode;
};

// New position of a node at the start of a pass.
template <int KIND>
__device__ inline double2 advance(const double2 x, const double2 v, const double2 f,
                                  const StepScal &S, const NetCtx &C, double2 &vh)
{
    double2 xn = x;
    if (KIND == kPassInit) {
        // affine_displace :533-536: x += xdisp (y - miny) / yrange, then y.
        if (S.pend_x != 0.0 || S.pend_y != 0.0) {
            xn.x = x.x + S.pend_x * (x.y - C.miny) / C.yrange;
            xn.y = x.y + S.pend_y * (x.y - C.miny) / C.yrange;
        }
        vh = make_double2(0.0, 0.0);
    } else if (KIND == kPassStep) {
        double2 vm;
        if (S.mix == kMixFire) {
            // :688-690: v <- v + (-alpha) v; v <- v + (alpha |v| / |F|) F
            vm.x = fma(S.cmix, f.x, fma(-S.alpha_old, v.x, v.x));
            vm.y = fma(S.cmix, f.y, fma(-S.alpha_old, v.y, v.y));
        } else if (S.mix == kMixZero) {
            vm = make_double2(0.0, 0.0);  // :711 memset after P <= 0
        } else {
            vm = v;
        }
        // :655-660: x += dt v; x += dt^2/(2m) F; v += dt/(2m) F
        xn.x = fma(S.dtsq, f.x, fma(S.dt, vm.x, x.x));
        xn.y = fma(S.dtsq, f.y, fma(S.dt, vm.y, x.y));
        vh.x = fma(S.dtm, f.x, vm.x);
        vh.y = fma(S.dtm, f.y, vm.y);
    }
    return xn;
}

// Force on own node k (local index) from its LDS-staged entries.
__device__ inline double2 node_force(const Lds &L, int k, int b_lo, int b_hi, int t_lo,
                                     int t_hi, const NetCtx &C, bool has_tkb)
{
    double fx = 0.0, fy = 0.0;
    const double2 self = L.pos[k];
    for (int q = b_lo; q < b_hi; q++) {
        const uint32_t code = L.bcode[q];
        const int nb = (int)(code & 0xFFFF);
        const int wx = wrap_of(code >> 16), wy = wrap_of(code >> 20);
        const double ox = C.W * wx + C.shift * wy, oy = C.H * wy;
        const double2 other = L.pos[nb];
        double rx, ry;
        const bool second = (code >> 24) & 1;
        if (!second) {
            rx = other.x + ox - self.x;
            ry = other.y + oy - self.y;
        } else {
            rx = self.x + ox - other.x;
            ry = self.y + oy - other.y;
        }
        const double d2 = fma(rx, rx, ry * ry);
        const double inv = 1.0 / sqrt(d2);
        const double s = C.ks * (1.0 - L.blen[q] * inv);
        const double bx = s * rx, by = s * ry;
        if (!second) {
            fx += bx;
            fy += by;
        } else {
            fx += -bx;
            fy += -by;
        }
    }
    for (int q = t_lo; q < t_hi; q++) {
        const uint64_t code = L.tcode[q];
        const int h = (int)(code & 0xFFFF), a = (int)((code >> 16) & 0xFFFF),
                  c = (int)((code >> 32) & 0xFFFF);
        const int wxa = wrap_of((uint32_t)(code >> 48)), wya = wrap_of((uint32_t)(code >> 52));
        const int wxb = wrap_of((uint32_t)(code >> 56)), wyb = wrap_of((uint32_t)(code >> 60));
        const double2 ph = L.pos[h], pa = L.pos[a], pc = L.pos[c];
        const double oxa = C.W * wxa + C.shift * wya, oya = C.H * wya;
        const double oxb = C.W * wxb + C.shift * wyb, oyb = C.H * wyb;
        const double ax = pa.x + oxa - ph.x, ay = pa.y + oya - ph.y;
        const double bx = pc.x + oxb - ph.x, by = pc.y + oyb - ph.y;
        const double ia = 1.0 / fma(ax, ax, ay * ay);
        const double ib = 1.0 / fma(bx, bx, by * by);
        const double cr = fma(ax, by, -(ay * bx));
        const double dt = fma(ax, bx, ay * by);
        // theta - pi of :450-453 as one atan2 (no cancellation near pi)
        const double dth = atan2(-cr, -dt);
        const double kk = has_tkb ? L.tkb[q] : C.kb;
        const double g = kk * dth;
        double f2x = g * ia;
        const double f2y = f2x * ax;
        f2x = f2x * -ay;
        double f3x = g * ib;
        const double f3y = -f3x * bx;
        f3x = f3x * by;
        if (k == h) {
            fx += -f2x - f3x;
            fy += -f2y - f3y;
        }
        if (k == a) {
            fx += f2x;
            fy += f2y;
        }
        if (k == c) {
            fx += f3x;
            fy += f3y;
        }
    }
    return make_double2(fx, fy);
}

// Energy of the elements own node k is responsible for: bonds where it is
// the first node, triples where it is the hinge.
__device__ inline void node_energy(const Lds &L, int k, int b_lo, int b_hi, int t_lo,
                                   int t_hi, const NetCtx &C, bool has_tkb, double &es,
                                   double &eb)
{
    const double2 self = L.pos[k];
    for (int q = b_lo; q < b_hi; q++) {
        const uint32_t code = L.bcode[q];
        if ((code >> 24) & 1) continue;
        const int nb = (int)(code & 0xFFFF);
        const int wx = wrap_of(code >> 16), wy = wrap_of(code >> 20);
        const double ox = C.W * wx + C.shift * wy, oy = C.H * wy;
        const double2 other = L.pos[nb];
        const double rx = other.x + ox - self.x, ry = other.y + oy - self.y;
        const double d = sqrt(fma(rx, rx, ry * ry));
        const double dl = d - L.blen[q];
        es += 0.5 * C.ks * dl * dl;
    }
    for (int q = t_lo; q < t_hi; q++) {
        const uint64_t code = L.tcode[q];
        const int h = (int)(code & 0xFFFF);
        if (h != k) continue;
        const int a = (int)((code >> 16) & 0xFFFF), c = (int)((code >> 32) & 0xFFFF);
        const int wxa = wrap_of((uint32_t)(code >> 48)), wya = wrap_of((uint32_t)(code >> 52));
        const int wxb = wrap_of((uint32_t)(code >> 56)), wyb = wrap_of((uint32_t)(code >> 60));
        const double2 pa = L.pos[a], pc = L.pos[c];
        const double oxa = C.W * wxa + C.shift * wya, oya = C.H * wya;
        const double oxb = C.W * wxb + C.shift * wyb, oyb = C.H * wyb;
        const double ax = pa.x + oxa - self.x, ay = pa.y + oya - self.y;
        const double bx = pc.x + oxb - self.x, by = pc.y + oyb - self.y;
        const double cr = fma(ax, by, -(ay * bx));
        const double dt = fma(ax, bx, ay * by);
        const double dth = atan2(-cr, -dt);
        const double kk = has_tkb ? L.tkb[q] : C.kb;
        eb += 0.5 * kk * dth * dth;
    }
}

// One pass over this member's tiles.  Init/step passes write x, v, F into
// buffer q = 1 - p and return (P, |v|^2, |F|^2, max |F_node|^2) partials;
// the energy pass reads buffer p and returns (E_stretch, E_bend, 0, 0).
template <int NPT, int KIND>
__device__ Red4 pass(const KArgs &A, const Team &T, const NetCtx &C, const StepScal &S,
                     int p, const Lds &L)
{
    Red4 acc = {0.0, 0.0, 0.0, 0.0};
    const int tid = threadIdx.x;
    const int q = p ^ 1;
    const double2 *__restrict__ X = A.x[p];
    const double2 *__restrict__ V = A.v[p];
    const double2 *__restrict__ F = A.f[p];
    const bool has_tkb = A.has_tkb != 0;
    for (int ti = T.member; ti < C.n_tiles; ti += T.g) {
        const TileDev td = A.tiles[C.tile_base + ti];
        // stage entries (coalesced) into LDS
        for (int i = tid; i < td.nb; i += kBlock) {
            L.bcode[i] = A.bcode[td.b0 + i];
            L.blen[i] = A.blen[td.b0 + i];
        }
        for (int i = tid; i < td.nt; i += kBlock) {
            L.tcode[i] = A.tcode[td.t0 + i];
            if (has_tkb) L.tkb[i] = A.tkb[td.t0 + i];
        }
        // positions: own nodes keep their half-kicked velocity in registers
        double2 vh[NPT];
#pragma unroll
        for (int j = 0; j < NPT; j++) {
            const int k = tid + j * kBlock;
            vh[j] = make_double2(0.0, 0.0);
            if (k < td.n_own) {
                const int gid = td.node0 + k;
                double2 xn;
                if (KIND == kPassEnergy) {
                    xn = X[gid];
                } else if (KIND == kPassInit) {
                    xn = advance<kPassInit>(X[gid], make_double2(0.0, 0.0), make_double2(0.0, 0.0),
                                            S, C, vh[j]);
                    A.x[q][gid] = xn;
                } else {
                    xn = advance<kPassStep>(X[gid], V[gid], F[gid], S, C, vh[j]);
                    A.x[q][gid] = xn;
                }
                L.pos[k] = xn;
            }
        }
        for (int k = td.n_own + tid; k < td.n_own + td.n_halo; k += kBlock) {
            const int gid = A.halo[td.halo_off + k - td.n_own];
            double2 dummy, xn;
            if (KIND == kPassEnergy)
                xn = X[gid];
            else if (KIND == kPassInit)
                xn = advance<kPassInit>(X[gid], make_double2(0.0, 0.0), make_double2(0.0, 0.0), S,
                                        C, dummy);
            else
                xn = advance<kPassStep>(X[gid], V[gid], F[gid], S, C, dummy);
            L.pos[k] = xn;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NPT; j++) {
            const int k = tid + j * kBlock;
            if (k < td.n_own) {
                const int gid = td.node0 + k;
                const int b_lo = A.brow[gid] - td.b0, b_hi = A.brow[gid + 1] - td.b0;
                const int t_lo = A.trow[gid] - td.t0, t_hi = A.trow[gid + 1] - td.t0;
                if (KIND == kPassEnergy) {
                    node_energy(L, k, b_lo, b_hi, t_lo, t_hi, C, has_tkb, acc.a, acc.b);
                } else {
                    double2 fn = node_force(L, k, b_lo, b_hi, t_lo, t_hi, C, has_tkb);
                    const unsigned char m = A.mask[gid];
                    if (!(m & 1)) fn.x = 0.0;
                    if (!(m & 2)) fn.y = 0.0;
                    double2 vn;
                    vn.x = fma(S.dtm, fn.x, vh[j].x);  // :677 second half kick
                    vn.y = fma(S.dtm, fn.y, vh[j].y);
                    if (KIND == kPassInit) vn = make_double2(0.0, 0.0);
                    A.f[q][gid] = fn;
                    A.v[q][gid] = vn;
                    acc.a = fma(vn.x, fn.x, acc.a);
                    acc.a = fma(vn.y, fn.y, acc.a);
                    acc.b = fma(vn.x, vn.x, acc.b);
                    acc.b = fma(vn.y, vn.y, acc.b);
                    const double f2 = fma(fn.x, fn.x, fn.y * fn.y);
                    acc.c += f2;
                    acc.m = fmax(acc.m, f2);
                }
            }
        }
        __syncthreads();
    }
    return acc;
}

template <int NPT>
__device__ void relax_network(const KArgs &A, Team &T, int net_id, const Lds &L,
                              double *s_red, int *s_flag)
{
    NetDev *nd = A.nets + net_id;
    NetCtx C;
    C.node_base = nd->node_base;
    C.n_nodes = nd->n_nodes;
    C.tile_base = nd->tile_base;
    C.n_tiles = nd->n_tiles;
    C.W = nd->width;
    C.H = nd->height;
    C.miny = nd->miny;
    C.yrange = nd->yrange;
    C.ks = nd->ks;
    C.kb = nd->kb;
    const double pend_x = nd->pend_x, pend_y = nd->pend_y;
    // the strain step advances the Lees-Edwards image shift before the pass
    C.shift = nd->le_shift + (C.H > 0 ? pend_x : 0.0);
    int p = nd->parity;
    const double a0 = nd->a0, fa = nd->fa, dtmax = nd->dtmax, finc = nd->finc,
                 fdec = nd->fdec, mass = nd->mass, fcut = nd->fcut;
    const int nmin = nd->nmin, max_steps = A.mode == kModeEval ? 0 : nd->max_steps;
    const int crit = nd->crit;

    StepScal S;
    S.dt = 0.0;
    S.dtsq = 0.0;
    S.dtm = 0.0;
    S.alpha_old = 0.0;
    S.cmix = 0.0;
    S.mix = kMixNone;
    S.pend_x = pend_x;
    S.pend_y = pend_y;

    int status = NR_NOT_CONVERGED;
    Red4 r;
    Red4 part = pass<NPT, kPassInit>(A, T, C, S, p, L);
    if (!reduce4(A, T, part, r, s_red, s_flag)) status = NR_ABORTED;
    p ^= 1;
    double ff = r.c, fmax2 = r.m;
    double fmag = sqrt(ff);
    const double sqrt_dof = sqrt((double)(2 * C.n_nodes));

    // run_fire_md :616-622
    double alpha = a0, dt = dtmax;
    double dtm = 0.5 * dt / mass, dtsq = 0.5 * dt * dt / mass;
    int since = 0, step = 0;
    S.mix = kMixNone;
    while (status != NR_ABORTED && step < max_steps) {
        step++;
        S.dt = dt;
        S.dtsq = dtsq;
        S.dtm = dtm;
        part = pass<NPT, kPassStep>(A, T, C, S, p, L);
        if (!reduce4(A, T, part, r, s_red, s_flag)) {
            status = NR_ABORTED;
            break;
        }
        p ^= 1;
        const double P = r.a, vv = r.b;
        ff = r.c;
        fmax2 = r.m;
        const double vmag = sqrt(vv);
        fmag = sqrt(ff);
        S.alpha_old = alpha;
        S.cmix = alpha * vmag / fmag;
        if (P > 0) {  // :694-703
            since++;
            if (since > nmin) {
                dt = fmin(dt * finc, dtmax);
                dtm = 0.5 * dt / mass;
                dtsq = 0.5 * dt * dt / mass;
                alpha *= fa;
            }
            S.mix = kMixFire;
        } else {      // :704-712
            since = 0;
            dt *= fdec;
            dtm = 0.5 * dt / mass;
            dtsq = 0.5 * dt * dt / mass;
            alpha = a0;
            S.mix = kMixZero;
        }
        const bool conv = crit == 1 ? sqrt(fmax2) < fcut : fmag / sqrt_dof < fcut;
        if (conv) break;
        if (!isfinite(ff)) break;
    }
    if (status != NR_ABORTED) {
        if (!isfinite(ff)) status = NR_NONFINITE;
        else {
            const double meas = crit == 1 ? sqrt(fmax2) : fmag / sqrt_dof;
            status = meas < fcut ? NR_CONVERGED : NR_NOT_CONVERGED;
        }
    }
    // energies at the final positions (:723-757)
    Red4 e = {0.0, 0.0, 0.0, 0.0};
    if (status != NR_ABORTED) {
        part = pass<NPT, kPassEnergy>(A, T, C, S, p, L);
        if (!reduce4(A, T, part, e, s_red, s_flag)) status = NR_ABORTED;
    }
    if (T.member == 0 && threadIdx.x == 0) {
        NetOut o;
        o.rms = fmag / sqrt_dof;
        o.fmax = sqrt(fmax2);
        o.es = C.ks > 0 ? e.a : 0.0;
        o.eb = C.kb > 0 ? e.b : 0.0;
        o.steps = step;
        o.status = status;
        o.fp64_steps = step;
        o.pad = 0;
        A.out[net_id] = o;
        nd->parity = p;
        nd->le_shift = C.shift;
        nd->pend_x = 0.0;
        nd->pend_y = 0.0;
    }
}

template <int NPT>
__global__ __launch_bounds__(kBlock) void relax_kernel(KArgs A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double s_red[4 * (kWaves + 1)];
    __shared__ int s_flag;
    __shared__ int s_job;
    Lds L;
    L.pos = reinterpret_cast<double2 *>(smem);
    L.blen = reinterpret_cast<double *>(L.pos + A.max_local);
    L.tcode = reinterpret_cast<uint64_t *>(L.blen + A.max_nb);
    L.tkb = reinterpret_cast<double *>(L.tcode + A.max_nt);
    L.bcode = reinterpret_cast<uint32_t *>(L.tkb + (A.has_tkb ? A.max_nt : 0));

    Team T;
    T.g = A.team;
    if (A.xcd_local) {
        const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
        T.id = xcd + 8 * (slot / T.g);
        T.member = slot % T.g;
    } else {
        T.id = blockIdx.x / T.g;
        T.member = blockIdx.x % T.g;
    }
    T.gen = 0;
    if (T.g > 1 && threadIdx.x == 0)
        T.gen = __hip_atomic_load(A.bar + T.id * 64 + 32, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);

    for (;;) {
        if (threadIdx.x == 0 && T.member == 0) {
            const int j = __hip_atomic_fetch_add(A.queue, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_job = j;
            if (T.g > 1)
                __hip_atomic_store(A.mailbox + T.id * 32, j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (T.g > 1) {
            if (!team_sync(A, T, &s_flag)) return;
            if (threadIdx.x == 0 && T.member != 0)
                s_job = __hip_atomic_load(A.mailbox + T.id * 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const int job = s_job;
        __syncthreads();
        if (job >= A.n_jobs) return;
        relax_network<NPT>(A, T, A.jobs[job], L, s_red, &s_flag);
        if (__hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    }
}

}  // namespace

int launch_relax(const KArgs &a, int npt, int grid, int block, size_t lds, bool cooperative,
                 void *stream)
{
    if (block != kBlock) return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    void *args[] = {const_cast<KArgs *>(&a)};
    const void *fn = nullptr;
    switch (npt) {
    case 1: fn = (const void *)relax_kernel<1>; break;
    case 2: fn = (const void *)relax_kernel<2>; break;
    case 4: fn = (const void *)relax_kernel<4>; break;
    case 8: fn = (const void *)relax_kernel<8>; break;
    default: return (int)hipErrorInvalidValue;
    }
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    if (cooperative)
        e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, (unsigned)lds, s);
    else
        e = hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, s);
    return (int)e;
}

int max_resident_blocks(int npt, int block, size_t lds, int *per_cu, int *n_cu)
{
    const void *fn = nullptr;
    switch (npt) {
    case 1: fn = (const void *)relax_kernel<1>; break;
    case 2: fn = (const void *)relax_kernel<2>; break;
    case 4: fn = (const void *)relax_kernel<4>; break;
    case 8: fn = (const void *)relax_kernel<8>; break;
    default: return (int)hipErrorInvalidValue;
    }
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    int nb = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, lds);
    if (e != hipSuccess) return (int)e;
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return (int)e;
    *per_cu = nb;
    *n_cu = prop.multiProcessorCount;
    return 0;
}

}  // namespace nr
