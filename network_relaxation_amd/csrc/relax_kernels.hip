// gfx950 kernels of the relax step: one persistent launch relaxes every
// network of a batch.  Workgroups form teams; a team takes networks from its
// size class's queue (longest expected relax first) and runs FIRE on each to
// convergence, each member owning a fixed set of spatial tiles; a team whose
// queue is empty splits into halves that serve the next smaller class.  Per
// FIRE step and tile:
//   1. the tile's topology blob is copied into LDS with 16-byte loads (once
//      per network when the blobs of all the member's tiles fit in LDS beside
//      the work area: they stay resident);
//   2. own + halo positions are advanced into LDS: the deferred FIRE mix,
//      drift and half kick of the previous step, or, on the first pass of a
//      relax, the affine / Lees-Edwards strain step;
//   3. element phase: each bond's and triple's force vectors are computed
//      once into LDS (one atan2 per triple);
//   4. gather: each own node adds its incident force vectors through its CSR
//      list in element order (no atomics, fixed order), half kick, partials.
// One exchange of tagged partial records per step combines the partials
// (P = v.F, |v|^2, |F|^2, max |F_node|^2) in member order, so every member
// takes the same FIRE decision.  References: run_fire_md large_strain_relax.cu:596-762,
// axial_forces :386-414, bending_forces :417-469, energies :472-523,
// affine_displace :527-537.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"

#pragma clang fp contract(off)

namespace nr {

namespace {

// kPassFuse starts the FP64 polish of a mixed-precision relax: x = FP64
// anchor + FP32 displacement, then forces as in kPassInit.
// kPassCgTrial is one line-search trial of nonlinear CG: direction
// d = f + beta d_old (first trial of an iteration; later trials reload it),
// trial point x + a d from the iteration's base point, forces there, and the
// reduction (F.d, F.F_base, |F|^2, max |F_node|^2).
enum PassKind { kPassInit = 0, kPassStep = 1, kPassEnergy = 2, kPassFuse = 3, kPassCgTrial = 4 };
enum MixKind { kMixNone = 0, kMixFire = 1, kMixZero = 2 };
// CG line search: at most this many secant trials, accepted when
// |phi'(a)| <= kCgEta |phi'(0)|.
constexpr int kCgLineSearch = 12;
constexpr double kCgEta = 1e-3;

struct Red4 {
    double a, b, c, m;  // sums a, b, c and max m
};

struct StepScal {
    double dt, dtsq, dtm;    // drift / kick of this pass
    double alpha_old, cmix;  // deferred FIRE mix of the previous step
    int mix;
    double pend_x, pend_y;   // affine step (init pass)
    int s;                   // parity of the FP32 state (mixed precision)
    double cg_a, cg_beta;    // CG trial: step along d, Fletcher-Reeves beta
    int cg_first;            // CG trial: form d (first trial) or reload it
    int pd;                  // CG: parity of the direction buffer d_old
};

struct NetCtx {
    int node_base, n_nodes, tile_base, n_tiles;
    double W, H, shift, miny, yrange, ks, kb;
};

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

// Data shared between the workgroups of a team (positions, velocities,
// forces, partial sums, the job mailbox) moves through agent-scope relaxed
// atomic loads and stores, which gfx950 performs coherently across XCDs
// without caching stale lines.  Before a workgroup arrives at a barrier every
// wave waits until its stores have completed, so the arrive atomic is ordered
// after them; the barrier therefore needs no L2 write-back or invalidate.
// A team of one workgroup uses ordinary cached accesses.
__device__ inline double2 ld_state(bool sh, const double2 *p)
{
    if (sh) {
        unsigned long long *q = reinterpret_cast<unsigned long long *>(const_cast<double2 *>(p));
        const unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_double2(__longlong_as_double((long long)a), __longlong_as_double((long long)b));
    }
    return *p;
}

// A shared double2 goes out as ONE 16-byte store with the agent-scope cache
// policy (sc1, what an agent-scope relaxed atomic store of 8 bytes uses):
// two 8-byte stores to the halves of 16-byte slots would write every 32-byte
// sector twice.  The store has no outputs; the drain before a record is
// published waits for it (vmcnt) like any other store.  A VALU write to the
// data registers of a store of more than 64 bits must wait one cycle after
// it; the compiler cannot see the store inside the asm (and has reused those
// registers in the very next instruction), so the wait is part of the asm.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ inline void st_state(bool sh, double2 *p, double2 v)
{
    if (sh) {
        const unsigned long long a = (unsigned long long)__double_as_longlong(v.x);
        const unsigned long long b = (unsigned long long)__double_as_longlong(v.y);
        u32x4 w;
        w.x = (unsigned)a;
        w.y = (unsigned)(a >> 32);
        w.z = (unsigned)b;
        w.w = (unsigned)(b >> 32);
        asm volatile("global_store_dwordx4 %0, %1, off sc1\n\ts_nop 1" : : "v"(p), "v"(w)
                     : "memory");
    } else {
        *p = v;
    }
}

// A shared double2 of the relax state comes in as ONE 16-byte load with the
// agent-scope cache policy (sc1), through a buffer descriptor over the
// network's node range (offsets relative to the network stay below 4 GB:
// networks have fewer than 2^27 nodes; reads past the range return zero).
constexpr int kBufWord3 = 0x00020000;  // raw buffer, gfx9 family
constexpr int kPolicySc1 = 16;         // cache policy bit sc1 (agent scope)
__device__ inline double2 ld_node(bool sh, const double2 *arr, const NetCtx &C, int gid)
{
    if (sh) {
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<double2 *>(arr + C.node_base), (short)0, C.n_nodes * 16, kBufWord3);
        const u32x4 w =
            __builtin_amdgcn_raw_buffer_load_b128(r, (gid - C.node_base) * 16, 0, kPolicySc1);
        return make_double2(
            __longlong_as_double((long long)(((unsigned long long)w.y << 32) | w.x)),
            __longlong_as_double((long long)(((unsigned long long)w.w << 32) | w.z)));
    }
    return arr[gid];
}

__device__ inline float2 ld_state32(bool sh, const float2 *p)
{
    if (sh) {
        const unsigned long long a = __hip_atomic_load(
            reinterpret_cast<unsigned long long *>(const_cast<float2 *>(p)), __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT);
        return make_float2(__uint_as_float((unsigned)(a & 0xFFFFFFFFull)),
                           __uint_as_float((unsigned)(a >> 32)));
    }
    return *p;
}

__device__ inline void st_state32(bool sh, float2 *p, float2 v)
{
    if (sh) {
        const unsigned long long a = (unsigned long long)__float_as_uint(v.x) |
                                     ((unsigned long long)__float_as_uint(v.y) << 32);
        __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), a, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *p = v;
    }
}

// Waits until every memory operation this wave issued has completed.
__device__ inline void drain_memory()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

struct Team {
    int id, member, g, part;
    unsigned long long gen;
    unsigned rtag;  // reductions so far in this launch (tag of the partial records)
};

// A member publishes its partial sums as 8 tagged 64-bit words (high half the
// step tag, low half 32 bits of a double), so a reader needs no barrier: it
// accepts a record when all 8 words carry the tag of the step it waits for.
__device__ inline void put_record(unsigned long long *w, unsigned tag, const Red4 &t)
{
    const unsigned long long hi = (unsigned long long)tag << 32;
    const double v[4] = {t.a, t.b, t.c, t.m};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v[i]);
        __hip_atomic_store(w + 2 * i, hi | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(w + 2 * i + 1, hi | (b & 0xFFFFFFFFull), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Bounded two-level team barrier, executed by thread 0 of each member after
// its workgroup drained its stores.  Arrival counters only grow: the member
// that brings its group's count to a multiple of the group size arrives at
// the root, and the one that completes the root publishes the next
// generation.  Returns false when the wait bound expired.
__device__ bool team_arrive(unsigned long long *bar, const Team &T, unsigned long long &gen,
                            long long timeout)
{
    const unsigned long long target = gen + 1;
    unsigned long long *root = bar, *genp = bar + 16;
    bool last;
    if (T.g <= kBarGroup) {
        last = (__hip_atomic_fetch_add(root, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1) %
                   (unsigned long long)T.g == 0;
    } else {
        const int grp = T.member / kBarGroup;
        const int n_groups = (T.g + kBarGroup - 1) / kBarGroup;
        const unsigned long long gsize = (unsigned long long)min(kBarGroup, T.g - grp * kBarGroup);
        unsigned long long *gc = bar + 32 + 16 * grp;
        last = (__hip_atomic_fetch_add(gc, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1) %
                   gsize == 0;
        if (last)
            last = (__hip_atomic_fetch_add(root, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1) %
                       (unsigned long long)n_groups == 0;
    }
    if (last) {
        __hip_atomic_store(genp, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(genp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != target) {
            __builtin_amdgcn_s_sleep(1);
            if (wall_clock64() - t0 > timeout) return false;
        }
    }
    drain_memory();
    gen = target;
    return true;
}

// Whole-workgroup team barrier; all threads return the same verdict.
__device__ bool team_sync(const KArgs &A, Team &T, int *s_flag)
{
    drain_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        bool ok = team_arrive(A.bar + (size_t)T.id * kBarStride, T, T.gen, A.timeout_ticks);
        if (!ok) __hip_atomic_store(A.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_flag = ok ? 1 : 0;
    }
    __syncthreads();
    return *s_flag != 0;
}

// First triple of thread tid in the element phase: triple chunk c (64
// consecutive triples) goes to wave (c + bond chunks) mod (B / 64).
template <int B>
__device__ inline int tri_first(int tid, int nb)
{
    constexpr int W = B / 64;
    const int wave = tid >> 6, nbc = (nb + 63) >> 6;
    return ((wave - nbc % W + W) % W) * 64 + (tid & 63);
}

// Block + team reduction of a Red4 in a fixed order (wave butterfly, waves
// in order; in a team, each member's block total is published as a tagged
// record and every member combines all records in the same fixed order).
// Every thread gets the result.
template <int B, int KS>
__device__ bool reduce4(const KArgs &A, Team &T, Red4 r, Red4 &out, double *s_red, int *s_flag)
{
    constexpr int W = B / 64;
    const bool team = T.g > 1;
    const unsigned tag = team ? ++T.rtag : 0u;
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
    // a member's record is published after every wave's state stores have
    // completed, so a member that reads it also sees that member's state
    if (team) drain_memory();
    __syncthreads();
    if (threadIdx.x == 0) {
        Red4 t = {0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < W; w++) {
            t.a += s_red[4 * w + 0];
            t.b += s_red[4 * w + 1];
            t.c += s_red[4 * w + 2];
            t.m = fmax(t.m, s_red[4 * w + 3]);
        }
        s_red[4 * W + 0] = t.a;
        s_red[4 * W + 1] = t.b;
        s_red[4 * W + 2] = t.c;
        s_red[4 * W + 3] = t.m;
        if (team) {
            *s_flag = 1;
            put_record(A.partials + (size_t)(T.part + T.member) * kPartWords + (tag & 1) * 8, tag, t);
        }
    }
    if (team) {
        __syncthreads();
        // Every member reads every member's record.  Four lanes share one:
        // lane q loads words 2q, 2q+1 (double q) in one 16-byte load, so a
        // wave load fetches 16 whole records and a spin costs one request
        // per record line.  Lane (r, q) polls records r, r + B/4, ... in two
        // groups of loads, and sums them in that order; lanes of equal q
        // then combine by butterfly, and waves in order.
        constexpr int R = B / 4;
        const int q = threadIdx.x & 3, rr = threadIdx.x >> 2;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            A.partials + (size_t)T.part * kPartWords, (short)0, T.g * kPartWords * 8, kBufWord3);
        const int boff = (int)(tag & 1) * 64 + q * 16;
        double got[KS];
        unsigned pending = 0;
#pragma unroll
        for (int k = 0; k < KS; k++) {
            got[k] = 0.0;
            if (rr + k * R < T.g) pending |= 1u << k;
        }
        bool ok = T.g <= KS * R;  // the host caps teams at max_team(B, NPT)
        if (!ok) pending = 0;
        const long long t0 = wall_clock64();
        int spins = 0;
        while (pending) {
#pragma unroll
            for (int k0 = 0; k0 < KS; k0 += 4) {
                u32x4 w[4];
#pragma unroll
                for (int k = k0; k < k0 + 4; k++)
                    if (pending & (1u << k))  // volatile (bit 31): reloaded every spin
                        w[k - k0] = __builtin_amdgcn_raw_buffer_load_b128(
                            rs, (rr + k * R) * (kPartWords * 8) + boff, 0,
                            kPolicySc1 | (int)0x80000000);
#pragma unroll
                for (int k = k0; k < k0 + 4; k++)
                    if ((pending & (1u << k)) && w[k - k0].y == tag && w[k - k0].w == tag) {
                        got[k] = __longlong_as_double(
                            (long long)(((unsigned long long)w[k - k0].x << 32) | w[k - k0].z));
                        pending &= ~(1u << k);
                    }
            }
            if (!pending) break;
            // the wait bound and the error flag are checked every 64 spins
            if ((++spins & 63) == 0 &&
                (wall_clock64() - t0 > A.timeout_ticks ||
                 __hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!ok) {
            __hip_atomic_store(A.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *s_flag = 0;
        }
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < KS; k++)
            if (rr + k * R < T.g) v = q == 3 ? fmax(v, got[k]) : v + got[k];
#pragma unroll
        for (int o = 4; o < 64; o <<= 1) {
            const double u = __shfl_xor(v, o, 64);
            v = q == 3 ? fmax(v, u) : v + u;
        }
        __syncthreads();
        if (lane < 4) s_red[4 * wave + lane] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            Red4 t = {0.0, 0.0, 0.0, 0.0};
            for (int w = 0; w < W; w++) {
                t.a += s_red[4 * w + 0];
                t.b += s_red[4 * w + 1];
                t.c += s_red[4 * w + 2];
                t.m = fmax(t.m, s_red[4 * w + 3]);
            }
            s_red[4 * W + 0] = t.a;
            s_red[4 * W + 1] = t.b;
            s_red[4 * W + 2] = t.c;
            s_red[4 * W + 3] = t.m;
        }
    }
    __syncthreads();
    out.a = s_red[4 * W + 0];
    out.b = s_red[4 * W + 1];
    out.c = s_red[4 * W + 2];
    out.m = s_red[4 * W + 3];
    const bool ok = !team || *s_flag != 0;
    __syncthreads();
    return ok;
}

// LDS regions: the tile blob, then positions and force vectors.
struct Lds {
    unsigned char *blob;
    double2 *pos;
    double2 *fv;
    TileDev *ctd;  // cached tile descriptors of this member, with the offset of
                   // each one's halo ids in chalo (lds_halo, -1: not cached) and
                   // of its resident blob from blob (lds_blob, -1: loaded per pass)
    int *chalo;    // halo id pool
};

// Descriptor of the kt-th tile of this member: from the LDS cache (every
// field moved to scalar registers) or, past the cache, from global memory.
// Built field by field so that it stays in registers.
__device__ inline TileDev tile_desc(const KArgs &A, const Lds &L0, int idx, int kt)
{
    TileDev u;
    if (kt < kTileCache) {
        const TileDev &t = L0.ctd[kt];
        u.net = __builtin_amdgcn_readfirstlane(t.net);
        u.node0 = __builtin_amdgcn_readfirstlane(t.node0);
        u.n_own = __builtin_amdgcn_readfirstlane(t.n_own);
        u.n_halo = __builtin_amdgcn_readfirstlane(t.n_halo);
        u.nb = __builtin_amdgcn_readfirstlane(t.nb);
        u.nt = __builtin_amdgcn_readfirstlane(t.nt);
        u.n_inc = __builtin_amdgcn_readfirstlane(t.n_inc);
        u.blob16 = __builtin_amdgcn_readfirstlane(t.blob16);
        u.blob_bytes = __builtin_amdgcn_readfirstlane(t.blob_bytes);
        u.lds_halo = __builtin_amdgcn_readfirstlane(t.lds_halo);
        u.lds_blob = __builtin_amdgcn_readfirstlane(t.lds_blob);
    } else {
        const TileDev *t = A.tiles + idx;
        u.net = t->net;
        u.node0 = t->node0;
        u.n_own = t->n_own;
        u.n_halo = t->n_halo;
        u.nb = t->nb;
        u.nt = t->nt;
        u.n_inc = t->n_inc;
        u.blob16 = t->blob16;
        u.blob_bytes = t->blob_bytes;
        u.lds_halo = -1;
        u.lds_blob = -1;
    }
    return u;
}

// Views into a tile blob held in LDS.
struct Blob {
    const int *halo;
    const uint32_t *bij;
    const uint8_t *bw;
    const double *bl;
    const uint64_t *tc;
    const double *tk;
    const uint16_t *ioff;
    const uint16_t *inc;
};

__device__ inline Blob blob_view(const unsigned char *base, const TileDev &td)
{
    const BlobLayout L = blob_layout(td.n_own, td.n_halo, td.nb, td.nt, td.n_inc);
    Blob b;
    b.halo = reinterpret_cast<const int *>(base + L.halo);
    b.bij = reinterpret_cast<const uint32_t *>(base + L.bij);
    b.bw = reinterpret_cast<const uint8_t *>(base + L.bw);
    b.bl = reinterpret_cast<const double *>(base + L.bl);
    b.tc = reinterpret_cast<const uint64_t *>(base + L.tc);
    b.tk = reinterpret_cast<const double *>(base + L.tk);
    b.ioff = reinterpret_cast<const uint16_t *>(base + L.ioff);
    b.inc = reinterpret_cast<const uint16_t *>(base + L.inc);
    return b;
}

// 16-byte blob loads a thread issues before the first of them is waited on.
constexpr int kBlobRound = 8;

// Copies a tile blob into LDS: 16-byte loads, four in flight per thread.
template <int B>
__device__ inline void stage_blob(const KArgs &A, const TileDev &td, unsigned char *dst_bytes)
{
    const uint4 *__restrict__ src = A.blob + td.blob16;
    uint4 *dst = reinterpret_cast<uint4 *>(dst_bytes);
    const int n16 = td.blob_bytes >> 4;
    for (int base = 0; base < n16; base += 4 * B) {
        // unconditional loads (index clamped into the blob) keep the four
        // loads in flight together and their results in registers
        const int i0 = base + (int)threadIdx.x, i1 = i0 + B, i2 = i0 + 2 * B, i3 = i0 + 3 * B;
        const uint4 r0 = src[min(i0, n16 - 1)];
        const uint4 r1 = src[min(i1, n16 - 1)];
        const uint4 r2 = src[min(i2, n16 - 1)];
        const uint4 r3 = src[min(i3, n16 - 1)];
        if (i0 < n16) dst[i0] = r0;
        if (i1 < n16) dst[i1] = r1;
        if (i2 < n16) dst[i2] = r2;
        if (i3 < n16) dst[i3] = r3;
    }
}

__device__ inline int nib(uint32_t v) { return (int)(v & 0xF) - kWrapBias; }

// Image offset of a neighbour with wraps (wx, wy): (wx W + wy shift, wy H).
__device__ inline void image(const NetCtx &C, int wx, int wy, double &ox, double &oy)
{
    ox = C.W * wx + C.shift * wy;
    oy = C.H * wy;
}

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
    } else if (KIND == kPassCgTrial) {
        // d = f + beta d_old, then x + a d (trial points from the base point)
        double2 d = v;
        if (S.cg_first) {
            d.x = fma(S.cg_beta, v.x, f.x);
            d.y = fma(S.cg_beta, v.y, f.y);
        }
        xn.x = fma(S.cg_a, d.x, x.x);
        xn.y = fma(S.cg_a, d.y, x.y);
        vh = d;
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

// 1/sqrt(x) and 1/x for the force terms: the hardware estimate refined by
// two Newton steps (y += y (1 - x y^2) / 2, y += y (1 - x y)), each of which
// squares the relative error of the estimate, so the result is within an ulp
// or so of the correctly rounded value; x = 0 gives a non-finite force as the
// correctly rounded forms do.
__device__ inline double rsqrt_newton(double x)
{
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const double r = fma(-(x * y), y, 1.0);
        y = fma(0.5 * y, r, y);
    }
    return y;
}

__device__ inline double rcp_newton(double x)
{
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int i = 0; i < 2; i++) y = fma(y, fma(-x, y, 1.0), y);
    return y;
}

// Bond vector r = x_j + image - x_i of bond record e (axial_forces :401-404,
// written as x_j + offset - x_i).
__device__ inline void bond_vec(const Blob &bb, const double2 *pos, const NetCtx &C, int e,
                                double &rx, double &ry)
{
    const uint32_t ij = bb.bij[e];
    const uint32_t w = bb.bw[e];
    const double2 pi = pos[ij & 0xFFFF], pj = pos[ij >> 16];
    double ox, oy;
    image(C, nib(w), nib(w >> 4), ox, oy);
    rx = pj.x + ox - pi.x;
    ry = pj.y + oy - pi.y;
}

// Arms of triple record t: a = x_a + img_a - x_h, b = x_b + img_b - x_h
// (bending_forces :447-450).
__device__ inline void tri_arms(const Blob &bb, const double2 *pos, const NetCtx &C, int t,
                                int &h, double &ax, double &ay, double &bx, double &by)
{
    const uint64_t code = bb.tc[t];
    h = (int)(code & 0xFFFF);
    const int a = (int)((code >> 16) & 0xFFFF), c = (int)((code >> 32) & 0xFFFF);
    const uint32_t w = (uint32_t)(code >> 48);
    const double2 ph = pos[h], pa = pos[a], pc = pos[c];
    double oxa, oya, oxb, oyb;
    image(C, nib(w), nib(w >> 4), oxa, oya);
    image(C, nib(w >> 8), nib(w >> 12), oxb, oyb);
    ax = pa.x + oxa - ph.x;
    ay = pa.y + oya - ph.y;
    bx = pc.x + oxb - ph.x;
    by = pc.y + oyb - ph.y;
}

// Force on own node k: its incident force vectors added in element order
// (the order of a sequential pass over the bond kernel, then the triple
// kernel).  Four entries are fetched before they are added.
__device__ inline double2 gather_force(const Blob &bb, const double2 *fv, int k)
{
    double fx = 0.0, fy = 0.0;
    const int lo = bb.ioff[k], hi = bb.ioff[k + 1];
    int q = lo;
    for (; q + 4 <= hi; q += 4) {
        double2 c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t e = bb.inc[q + u];
            const double2 v = fv[e & kMaxForceVectors];
            c[u] = (e & kIncNegate) ? make_double2(-v.x, -v.y) : v;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            fx += c[u].x;
            fy += c[u].y;
        }
    }
    for (; q < hi; q++) {
        const uint32_t e = bb.inc[q];
        const double2 v = fv[e & kMaxForceVectors];
        if (e & kIncNegate) {
            fx += -v.x;
            fy += -v.y;
        } else {
            fx += v.x;
            fy += v.y;
        }
    }
    return make_double2(fx, fy);
}

// FIRE scalar state and its per-step update (run_fire_md :616-622 and
// :679-712): the mix of the step just taken is deferred to the next pass's
// position update (S.alpha_old, S.cmix, S.mix).
struct Fire {
    const NetDev *nd;  // FIRE constants, read where used (scalar cache)
    double alpha, dt;
    int since;

    __device__ void start()
    {
        alpha = nd->a0;
        dt = nd->dtmax;
        since = 0;
    }
    __device__ void load(StepScal &S) const
    {
        const double mass = nd->mass;
        S.dt = dt;
        S.dtsq = 0.5 * dt * dt / mass;
        S.dtm = 0.5 * dt / mass;
    }
    __device__ void update(const Red4 &r, StepScal &S)
    {
        const double P = r.a, vmag = sqrt(r.b), fmag = sqrt(r.c);
        S.alpha_old = alpha;
        S.cmix = alpha * vmag / fmag;
        if (P > 0) {  // :694-703
            since++;
            if (since > nd->nmin) {
                dt = fmin(dt * nd->finc, nd->dtmax);
                alpha *= nd->fa;
            }
            S.mix = kMixFire;
        } else {  // :704-712
            since = 0;
            dt *= nd->fdec;
            alpha = nd->a0;
            S.mix = kMixZero;
        }
    }
};

// x = FP64 anchor + FP32 displacement (start of the FP64 polish).
__device__ inline double2 fuse_position(bool sh, const KArgs &A, const StepScal &S,
                                        const double2 *X, int gid)
{
    const double2 a = ld_state(sh, X + gid);
    const float2 u = ld_state32(sh, A.u32[S.s] + gid);
    return make_double2(a.x + (double)u.x, a.y + (double)u.y);
}

// ---- FP32 stage of a mixed-precision relax ------------------------------
// State: FP64 anchor a (positions after the strain step, fixed during the
// stage) and FP32 displacement u, velocity v and force F.  Every geometric
// term is split into an FP64 anchor part and FP32 displacement corrections,
// so that FP32 rounding is relative to the displacement, not to the bond:
//   bond:   d^2 - L^2 = (|a|^2 - L^2) + 2 a.du + du.du,
//           f = ks (d^2 - L^2) / (d (d + L)) r      [= ks (1 - L/d) r]
//   triple: cross = a_a x a_b + a_a x du_b + du_a x a_b + du_a x du_b,
//           dot likewise, theta - pi = atan2(-cross, -dot).
enum Pass32Kind { kPass32Init = 0, kPass32Step = 1 };

struct Lds32 {
    unsigned char *blob;
    double2 *anc;  // FP64 anchors (own + halo)
    float2 *u;     // FP32 displacements (own + halo)
    float2 *fv;    // FP32 force vectors
};

template <int KIND>
__device__ inline float2 advance32(const float2 u, const float2 v, const float2 f,
                                   const StepScal &S, float2 &vh)
{
    if (KIND == kPass32Init) {
        vh = make_float2(0.0f, 0.0f);
        return make_float2(0.0f, 0.0f);
    }
    float2 vm;
    if (S.mix == kMixFire) {
        const float na = (float)-S.alpha_old, cm = (float)S.cmix;
        vm.x = fmaf(cm, f.x, fmaf(na, v.x, v.x));
        vm.y = fmaf(cm, f.y, fmaf(na, v.y, v.y));
    } else if (S.mix == kMixZero) {
        vm = make_float2(0.0f, 0.0f);
    } else {
        vm = v;
    }
    const float dt = (float)S.dt, dtsq = (float)S.dtsq, dtm = (float)S.dtm;
    float2 un;
    un.x = fmaf(dtsq, f.x, fmaf(dt, vm.x, u.x));
    un.y = fmaf(dtsq, f.y, fmaf(dt, vm.y, u.y));
    vh.x = fmaf(dtm, f.x, vm.x);
    vh.y = fmaf(dtm, f.y, vm.y);
    return un;
}

// One FP32-stage pass over this member's tiles.  Reads the anchor from x[pa]
// (the init pass first writes it there from x[p] and the strain step) and the
// FP32 state from parity S.s; writes the FP32 state to parity S.s ^ 1.
template <int B, int NPT, int KIND>
__device__ Red4 pass32(const KArgs &A, const Team &T, const NetCtx &C, const StepScal &S,
                       int p, int pa, const Lds &L0)
{
    Red4 acc = {0.0, 0.0, 0.0, 0.0};
    const int tid = threadIdx.x;
    const int s = S.s, s1 = S.s ^ 1;
    const bool sh = T.g > 1;
    const float2 zf = make_float2(0.0f, 0.0f);
    const float ks = (float)C.ks;
    for (int ti = T.member, kt = 0; ti < C.n_tiles; ti += T.g, kt++) {
        const bool cached = kt < kTileCache;
        const TileDev td = tile_desc(A, L0, C.tile_base + ti, kt);
        const int hoff = td.lds_halo, roff = td.lds_blob;
        const bool resident = roff >= 0;
        const int local = td.n_own + td.n_halo;
        // a resident blob stays at the top of the tile area (with two nodes
        // per thread: where a loaded one goes); a loaded one goes to its
        // start, the anchors, displacements and force vectors after it
        const bool top = NPT == 1 && resident;
        Lds32 L;
        L.blob = L0.blob + (top ? roff : 0);
        L.anc = reinterpret_cast<double2 *>(L0.blob + (top ? 0 : align16(td.blob_bytes)));
        L.u = reinterpret_cast<float2 *>(L.anc + local);
        L.fv = L.u + local;
        // halo ids already in LDS (resident blob or the member's cache): the
        // blob, own-node and halo-node loads then go out together
        const int *hid = resident ? reinterpret_cast<const int *>(L.blob)
                         : cached && hoff >= 0 ? L0.chalo + hoff
                                               : nullptr;
        const uint4 *__restrict__ bsrc = A.blob + td.blob16;
        uint4 *bdst = reinterpret_cast<uint4 *>(L.blob);
        const int n16 = resident ? 0 : td.blob_bytes >> 4;
        uint4 r0, r1, r2, r3, r4, r5, r6, r7;
        if (!resident) {
            const int last = n16 - 1;
            r0 = bsrc[min(tid, last)];
            r1 = bsrc[min(tid + B, last)];
            r2 = bsrc[min(tid + 2 * B, last)];
            r3 = bsrc[min(tid + 3 * B, last)];
            r4 = bsrc[min(tid + 4 * B, last)];
            r5 = bsrc[min(tid + 5 * B, last)];
            r6 = bsrc[min(tid + 6 * B, last)];
            r7 = bsrc[min(tid + 7 * B, last)];
        }
        // own node: anchor (or, on the init pass, the position before the
        // strain step) and FP32 state
        const int k = tid;
        double2 xo = make_double2(0.0, 0.0);
        float2 uo = zf, vo = zf, fo = zf;
        unsigned char mk = 0;
        if (k < td.n_own) {
            const int gid = td.node0 + k;
            xo = ld_node(sh, KIND == kPass32Init ? A.x[p] : A.x[pa], C, gid);
            if (KIND != kPass32Init) {
                uo = ld_state32(sh, A.u32[s] + gid);
                vo = ld_state32(sh, A.v32[s] + gid);
                fo = ld_state32(sh, A.f32[s] + gid);
            }
            mk = A.mask[gid];
        }
        // first halo node of this thread (the rest, if any, below)
        const bool hfirst = hid != nullptr && tid < td.n_halo;
        double2 hx = make_double2(0.0, 0.0);
        float2 hu = zf, hv = zf, hf = zf;
        if (hfirst) {
            const int gid = hid[tid];
            hx = ld_node(sh, KIND == kPass32Init ? A.x[p] : A.x[pa], C, gid);
            if (KIND != kPass32Init) {
                hu = ld_state32(sh, A.u32[s] + gid);
                hv = ld_state32(sh, A.v32[s] + gid);
                hf = ld_state32(sh, A.f32[s] + gid);
            }
        }
        if (!resident) {
            if (tid < n16) bdst[tid] = r0;
            if (tid + B < n16) bdst[tid + B] = r1;
            if (tid + 2 * B < n16) bdst[tid + 2 * B] = r2;
            if (tid + 3 * B < n16) bdst[tid + 3 * B] = r3;
            if (tid + 4 * B < n16) bdst[tid + 4 * B] = r4;
            if (tid + 5 * B < n16) bdst[tid + 5 * B] = r5;
            if (tid + 6 * B < n16) bdst[tid + 6 * B] = r6;
            if (tid + 7 * B < n16) bdst[tid + 7 * B] = r7;
            for (int i = kBlobRound * B + tid; i < n16; i += B) bdst[i] = bsrc[i];
        }
        float2 vh = zf;
        if (k < td.n_own) {
            const int gid = td.node0 + k;
            double2 a;
            float2 un;
            if (KIND == kPass32Init) {
                double2 dummy;
                a = advance<kPassInit>(xo, make_double2(0.0, 0.0), make_double2(0.0, 0.0), S, C,
                                       dummy);
                st_state(sh, A.x[pa] + gid, a);
                un = zf;
            } else {
                a = xo;
                un = advance32<kPass32Step>(uo, vo, fo, S, vh);
            }
            st_state32(sh, A.u32[s1] + gid, un);
            L.anc[k] = a;
            L.u[k] = un;
        }
        if (hfirst) {
            double2 a;
            float2 un, dummy;
            if (KIND == kPass32Init) {
                double2 d2;
                a = advance<kPassInit>(hx, make_double2(0.0, 0.0), make_double2(0.0, 0.0), S, C,
                                       d2);
                un = zf;
            } else {
                a = hx;
                un = advance32<kPass32Step>(hu, hv, hf, S, dummy);
            }
            L.anc[td.n_own + tid] = a;
            L.u[td.n_own + tid] = un;
        }
        if (hid == nullptr) __syncthreads();  // halo ids come from the blob
        const Blob bb = blob_view(L.blob, td);
        for (int kk = td.n_own + (hid != nullptr ? B : 0) + tid; kk < local; kk += B) {
            const int gid = hid != nullptr ? hid[kk - td.n_own] : bb.halo[kk - td.n_own];
            double2 a;
            float2 un, dummy;
            if (KIND == kPass32Init) {
                double2 d2;
                a = advance<kPassInit>(ld_node(sh, A.x[p], C, gid), make_double2(0.0, 0.0),
                                       make_double2(0.0, 0.0), S, C, d2);
                un = zf;
            } else {
                a = ld_node(sh, A.x[pa], C, gid);
                un = advance32<kPass32Step>(ld_state32(sh, A.u32[s] + gid),
                                            ld_state32(sh, A.v32[s] + gid),
                                            ld_state32(sh, A.f32[s] + gid), S, dummy);
            }
            L.anc[kk] = a;
            L.u[kk] = un;
        }
        __syncthreads();
        for (int e = tid; e < td.nb; e += B) {
            const uint32_t ij = bb.bij[e];
            const uint32_t w = bb.bw[e];
            const int i = (int)(ij & 0xFFFF), j = (int)(ij >> 16);
            double ox, oy;
            image(C, nib(w), nib(w >> 4), ox, oy);
            const double2 ai = L.anc[i], aj = L.anc[j];
            const double axd = aj.x + ox - ai.x, ayd = aj.y + oy - ai.y;
            const double L64 = bb.bl[e];
            const float s0 = (float)(fma(axd, axd, ayd * ayd) - L64 * L64);
            const float2 ui = L.u[i], uj = L.u[j];
            const float dx = uj.x - ui.x, dy = uj.y - ui.y;
            const float ax = (float)axd, ay = (float)ayd;
            const float rx = ax + dx, ry = ay + dy;
            const float d2ml2 = s0 + fmaf(2.0f * ax + dx, dx, (2.0f * ay + dy) * dy);
            const float d = sqrtf(fmaf(rx, rx, ry * ry));
            const float coef = ks * d2ml2 / (d * (d + (float)L64));
            L.fv[e] = make_float2(coef * rx, coef * ry);
        }
        // triples are dealt to the waves starting where the bonds' 64-element
        // chunks ended, so that every wave gets the same number of chunks
        for (int t = tri_first<B>(tid, td.nb); t < td.nt; t += B) {
            const uint64_t code = bb.tc[t];
            const int h = (int)(code & 0xFFFF), a = (int)((code >> 16) & 0xFFFF),
                      c = (int)((code >> 32) & 0xFFFF);
            const uint32_t w = (uint32_t)(code >> 48);
            double oxa, oya, oxb, oyb;
            image(C, nib(w), nib(w >> 4), oxa, oya);
            image(C, nib(w >> 8), nib(w >> 12), oxb, oyb);
            const double2 ah = L.anc[h], aa = L.anc[a], ac = L.anc[c];
            const double pax = aa.x + oxa - ah.x, pay = aa.y + oya - ah.y;
            const double pbx = ac.x + oxb - ah.x, pby = ac.y + oyb - ah.y;
            const float cr0 = (float)fma(pax, pby, -(pay * pbx));
            const float dt0 = (float)fma(pax, pbx, pay * pby);
            const float2 uh = L.u[h], ua = L.u[a], uc = L.u[c];
            const float dax = ua.x - uh.x, day = ua.y - uh.y;
            const float dbx = uc.x - uh.x, dby = uc.y - uh.y;
            const float fax = (float)pax, fay = (float)pay, fbx = (float)pbx, fby = (float)pby;
            const float cr = cr0 + ((fax * dby - fay * dbx) + (dax * fby - day * fbx) +
                                    (dax * dby - day * dbx));
            const float dt = dt0 + ((fax * dbx + fay * dby) + (dax * fbx + day * fby) +
                                    (dax * dbx + day * dby));
            const float dth = atan2f(-cr, -dt);
            const float rax = fax + dax, ray = fay + day, rbx = fbx + dbx, rby = fby + dby;
            const float g = (float)bb.tk[t] * dth;
            const float ia = g / fmaf(rax, rax, ray * ray);
            const float ib = g / fmaf(rbx, rbx, rby * rby);
            const float2 f2 = make_float2(ia * -ray, ia * rax);
            const float2 f3 = make_float2(ib * rby, -ib * rbx);
            L.fv[td.nb + t] = f2;
            L.fv[td.nb + td.nt + t] = f3;
            L.fv[td.nb + 2 * td.nt + t] = make_float2(-f2.x - f3.x, -f2.y - f3.y);
        }
        __syncthreads();
        if (k < td.n_own) {
            const int gid = td.node0 + k;
            double fx = 0.0, fy = 0.0;
            const int lo = bb.ioff[k], hi = bb.ioff[k + 1];
            for (int q = lo; q < hi; q++) {
                const uint32_t e = bb.inc[q];
                const float2 v = L.fv[e & kMaxForceVectors];
                if (e & kIncNegate) {
                    fx -= (double)v.x;
                    fy -= (double)v.y;
                } else {
                    fx += (double)v.x;
                    fy += (double)v.y;
                }
            }
            if (!(mk & 1)) fx = 0.0;
            if (!(mk & 2)) fy = 0.0;
            const float2 fn = make_float2((float)fx, (float)fy);
            float2 vn;
            if (KIND == kPass32Init) {
                vn = zf;
            } else {
                const float dtm = (float)S.dtm;
                vn = make_float2(fmaf(dtm, fn.x, vh.x), fmaf(dtm, fn.y, vh.y));
            }
            st_state32(sh, A.f32[s1] + gid, fn);
            st_state32(sh, A.v32[s1] + gid, vn);
            const double vx = vn.x, vy = vn.y, gx = fn.x, gy = fn.y;
            acc.a = fma(vx, gx, acc.a);
            acc.a = fma(vy, gy, acc.a);
            acc.b = fma(vx, vx, acc.b);
            acc.b = fma(vy, vy, acc.b);
            const double f2 = fma(gx, gx, gy * gy);
            acc.c += f2;
            acc.m = fmax(acc.m, f2);
        }
        __syncthreads();
    }
    return acc;
}

// One pass over this member's tiles.  Init/step passes write x, v, F into
// buffer q = 1 - p and return (P, |v|^2, |F|^2, max |F_node|^2) partials;
// the energy pass reads buffer p and returns (E_stretch, E_bend, 0, 0).
template <int B, int NPT, int KIND>
__device__ Red4 pass(const KArgs &A, const Team &T, const NetCtx &C, const StepScal &S, int p,
                     const Lds &L0)
{
    Red4 acc = {0.0, 0.0, 0.0, 0.0};
    const int tid = threadIdx.x;
    const int q = p ^ 1;
    const double2 *__restrict__ X = A.x[p];
    // CG keeps its direction in the velocity buffers: d_old in v[pd], the
    // iteration's d (written by its first trial) in v[pd ^ 1]
    const double2 *__restrict__ V =
        KIND == kPassCgTrial ? A.v[S.cg_first ? S.pd : S.pd ^ 1] : A.v[p];
    const double2 *__restrict__ F = A.f[p];
    const double2 zero = make_double2(0.0, 0.0);
    const bool sh = T.g > 1;
    for (int ti = T.member, kt = 0; ti < C.n_tiles; ti += T.g, kt++) {
        const bool cached = kt < kTileCache;
        const TileDev td = tile_desc(A, L0, C.tile_base + ti, kt);
        const int hoff = td.lds_halo, roff = td.lds_blob;
        const bool resident = roff >= 0;
        // a resident blob stays at the top of the tile area (with two nodes
        // per thread: where a loaded one goes); a loaded one goes to its
        // start, the positions and force vectors after it
        const bool top = NPT == 1 && resident;
        Lds L = L0;
        L.blob = L0.blob + (top ? roff : 0);
        L.pos = reinterpret_cast<double2 *>(L0.blob + (top ? 0 : align16(td.blob_bytes)));
        L.fv = L.pos + td.n_own + td.n_halo;
        // halo ids already in LDS (resident blob or the member's cache): the
        // blob, own-node and halo-node loads then go out together
        const int *hid = resident ? reinterpret_cast<const int *>(L.blob)
                         : cached && hoff >= 0 ? L0.chalo + hoff
                                               : nullptr;
        // The blob's first kBlobRound x 16 B per thread and the own nodes'
        // state are loaded together; the LDS writes then wait only for the
        // loads they use.
        const uint4 *__restrict__ bsrc = A.blob + td.blob16;
        uint4 *bdst = reinterpret_cast<uint4 *>(L.blob);
        const int n16 = resident ? 0 : td.blob_bytes >> 4;
        uint4 r0, r1, r2, r3, r4, r5, r6, r7;
        if (!resident) {
            const int last = n16 - 1;
            r0 = bsrc[min(tid, last)];
            r1 = bsrc[min(tid + B, last)];
            r2 = bsrc[min(tid + 2 * B, last)];
            r3 = bsrc[min(tid + 3 * B, last)];
            r4 = bsrc[min(tid + 4 * B, last)];
            r5 = bsrc[min(tid + 5 * B, last)];
            r6 = bsrc[min(tid + 6 * B, last)];
            r7 = bsrc[min(tid + 7 * B, last)];
        }
        double2 xo[NPT], vo[NPT], fo[NPT];
        unsigned char mk[NPT];
#pragma unroll
        for (int j = 0; j < NPT; j++) {
            const int k = tid + j * B;
            xo[j] = vo[j] = fo[j] = zero;
            mk[j] = 0;
            if (k < td.n_own) {
                const int gid = td.node0 + k;
                if (KIND == kPassFuse) {
                    xo[j] = fuse_position(sh, A, S, X, gid);
                    const float2 v32 = ld_state32(sh, A.v32[S.s] + gid);
                    vo[j] = make_double2((double)v32.x, (double)v32.y);
                } else {
                    xo[j] = ld_node(sh, X, C, gid);
                }
                if (KIND == kPassStep || KIND == kPassCgTrial) {
                    vo[j] = ld_node(sh, V, C, gid);
                    fo[j] = ld_node(sh, F, C, gid);
                }
                if (KIND != kPassEnergy) mk[j] = A.mask[gid];
            }
        }
        // first halo node of this thread (the rest, if any, below)
        const bool hfirst = hid != nullptr && tid < td.n_halo;
        double2 hx = zero, hv = zero, hf = zero;
        if (hfirst) {
            const int gid = hid[tid];
            if (KIND == kPassFuse) {
                hx = fuse_position(sh, A, S, X, gid);
            } else {
                hx = ld_node(sh, X, C, gid);
                if (KIND == kPassStep || KIND == kPassCgTrial) {
                    hv = ld_node(sh, V, C, gid);
                    hf = ld_node(sh, F, C, gid);
                }
            }
        }
        if (!resident) {
            if (tid < n16) bdst[tid] = r0;
            if (tid + B < n16) bdst[tid + B] = r1;
            if (tid + 2 * B < n16) bdst[tid + 2 * B] = r2;
            if (tid + 3 * B < n16) bdst[tid + 3 * B] = r3;
            if (tid + 4 * B < n16) bdst[tid + 4 * B] = r4;
            if (tid + 5 * B < n16) bdst[tid + 5 * B] = r5;
            if (tid + 6 * B < n16) bdst[tid + 6 * B] = r6;
            if (tid + 7 * B < n16) bdst[tid + 7 * B] = r7;
            for (int i = kBlobRound * B + tid; i < n16; i += B) bdst[i] = bsrc[i];
        }
        // positions: own nodes keep their half-kicked velocity and clamp mask
        // in registers
        double2 vh[NPT];
#pragma unroll
        for (int j = 0; j < NPT; j++) {
            const int k = tid + j * B;
            vh[j] = zero;
            if (k < td.n_own) {
                const int gid = td.node0 + k;
                double2 xn;
                if (KIND == kPassEnergy) {
                    xn = xo[j];
                } else if (KIND == kPassInit) {
                    xn = advance<kPassInit>(xo[j], zero, zero, S, C, vh[j]);
                    st_state(sh, A.x[q] + gid, xn);
                } else if (KIND == kPassFuse) {
                    xn = xo[j];
                    vh[j] = vo[j];
                    st_state(sh, A.x[q] + gid, xn);
                } else if (KIND == kPassCgTrial) {
                    xn = advance<kPassCgTrial>(xo[j], vo[j], fo[j], S, C, vh[j]);
                    st_state(sh, A.x[q] + gid, xn);
                    if (S.cg_first) st_state(sh, A.v[S.pd ^ 1] + gid, vh[j]);
                } else {
                    xn = advance<kPassStep>(xo[j], vo[j], fo[j], S, C, vh[j]);
                    st_state(sh, A.x[q] + gid, xn);
                }
                L.pos[k] = xn;
            }
        }
        // halo nodes: advanced here as their owners advance them
        if (hfirst) {
            double2 dummy, xn;
            if (KIND == kPassEnergy || KIND == kPassFuse)
                xn = hx;
            else if (KIND == kPassInit)
                xn = advance<kPassInit>(hx, zero, zero, S, C, dummy);
            else
                xn = advance<KIND>(hx, hv, hf, S, C, dummy);
            L.pos[td.n_own + tid] = xn;
        }
        if (hid == nullptr) __syncthreads();  // halo ids come from the blob
        const Blob bb = blob_view(L.blob, td);
        for (int k = td.n_own + (hid != nullptr ? B : 0) + tid; k < td.n_own + td.n_halo; k += B) {
            const int gid = hid != nullptr ? hid[k - td.n_own] : bb.halo[k - td.n_own];
            double2 dummy, xn;
            if (KIND == kPassEnergy)
                xn = ld_node(sh, X, C, gid);
            else if (KIND == kPassInit)
                xn = advance<kPassInit>(ld_node(sh, X, C, gid), zero, zero, S, C, dummy);
            else if (KIND == kPassFuse)
                xn = fuse_position(sh, A, S, X, gid);
            else
                xn = advance<KIND>(ld_node(sh, X, C, gid), ld_node(sh, V, C, gid),
                                   ld_node(sh, F, C, gid), S, C, dummy);
            L.pos[k] = xn;
        }
        __syncthreads();
        // element phase
        for (int e = tid; e < td.nb; e += B) {
            double rx, ry;
            bond_vec(bb, L.pos, C, e, rx, ry);
            if (KIND == kPassEnergy) {
                if ((int)(bb.bij[e] & 0xFFFF) < td.n_own) {
                    const double d = sqrt(fma(rx, rx, ry * ry));
                    const double dl = d - bb.bl[e];
                    acc.a += 0.5 * C.ks * dl * dl;
                    // shear virial dE/dgamma: ks (1 - L/|r|) r_x r_y
                    const double s = C.ks * (1.0 - bb.bl[e] * (1.0 / d));
                    acc.c += s * rx * ry;
                }
            } else {
                // :401-404: f = ks (1 - L/|r|) r on the first node
                const double d2 = fma(rx, rx, ry * ry);
                const double inv = rsqrt_newton(d2);
                const double s = C.ks * (1.0 - bb.bl[e] * inv);
                L.fv[e] = make_double2(s * rx, s * ry);
            }
        }
        // triples are dealt to the waves starting where the bonds' 64-element
        // chunks ended, so that every wave gets the same number of chunks
        for (int t = tri_first<B>(tid, td.nb); t < td.nt; t += B) {
            int h;
            double ax, ay, bx, by;
            tri_arms(bb, L.pos, C, t, h, ax, ay, bx, by);
            const double cr = fma(ax, by, -(ay * bx));
            const double dt = fma(ax, bx, ay * by);
            // theta - pi of :450-453 as one atan2 (no cancellation near pi)
            const double dth = atan2(-cr, -dt);
            const double kk = bb.tk[t];
            if (KIND == kPassEnergy) {
                if (h < td.n_own) {
                    acc.b += 0.5 * kk * dth * dth;
                    // shear virial: -(F_a,x a_y + F_b,x b_y)
                    const double g = kk * dth;
                    const double f2x = g * (1.0 / fma(ax, ax, ay * ay)) * -ay;
                    const double f3x = g * (1.0 / fma(bx, bx, by * by)) * by;
                    acc.c -= f2x * ay + f3x * by;
                }
            } else {
                // :453-468 force on end a, end b and the hinge
                const double ia = rcp_newton(fma(ax, ax, ay * ay));
                const double ib = rcp_newton(fma(bx, bx, by * by));
                const double g = kk * dth;
                double f2x = g * ia;
                const double f2y = f2x * ax;
                f2x = f2x * -ay;
                double f3x = g * ib;
                const double f3y = -f3x * bx;
                f3x = f3x * by;
                L.fv[td.nb + t] = make_double2(f2x, f2y);
                L.fv[td.nb + td.nt + t] = make_double2(f3x, f3y);
                L.fv[td.nb + 2 * td.nt + t] = make_double2(-f2x - f3x, -f2y - f3y);
            }
        }
        if (KIND != kPassEnergy) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < NPT; j++) {
                const int k = tid + j * B;
                if (k < td.n_own) {
                    const int gid = td.node0 + k;
                    double2 fn = gather_force(bb, L.fv, k);
                    if (!(mk[j] & 1)) fn.x = 0.0;
                    if (!(mk[j] & 2)) fn.y = 0.0;
                    st_state(sh, A.f[q] + gid, fn);
                    if (KIND == kPassCgTrial) {
                        // F . d along the trial direction, and F . F_base
                        // (the Polak-Ribiere numerator's cross term)
                        acc.a = fma(fn.x, vh[j].x, acc.a);
                        acc.a = fma(fn.y, vh[j].y, acc.a);
                        acc.b = fma(fn.x, fo[j].x, acc.b);
                        acc.b = fma(fn.y, fo[j].y, acc.b);
                    } else {
                        double2 vn;
                        vn.x = fma(S.dtm, fn.x, vh[j].x);  // :677 second half kick
                        vn.y = fma(S.dtm, fn.y, vh[j].y);
                        if (KIND == kPassInit) vn = zero;
                        if (KIND == kPassFuse) vn = vh[j];  // the FP32 stage's velocity
                        st_state(sh, A.v[q] + gid, vn);
                        acc.a = fma(vn.x, fn.x, acc.a);
                        acc.a = fma(vn.y, fn.y, acc.a);
                        acc.b = fma(vn.x, vn.x, acc.b);
                        acc.b = fma(vn.y, vn.y, acc.b);
                    }
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

// CG selects the method compiled in: nonlinear CG, or FIRE (with the mixed
// precision stage).  Each is its own kernel so that neither pays the other's
// registers.
template <int B, int NPT, bool CG>
__device__ void relax_network(const KArgs &A, Team &T, int net_id, const Lds &L, double *s_red,
                              int *s_flag)
{
    NetDev *nd = A.nets + net_id;
    NetRun *rs = A.run + net_id;
    const bool lead = T.member == 0 && threadIdx.x == 0;  // writes the job's records
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
    double shift = nd->le_shift;
    int p = nd->parity;
    const int max_steps = A.mode == kModeEval ? 0 : nd->max_steps;
    const int slice = rs->slice > 0 ? rs->slice : A.slice_steps;
    const int budget = slice > 0 ? slice : 0x7fffffff;
    int used = 0;
    // resumable state kept in registers; the rest of NetRun is read where needed
    int phase = rs->phase, inc_next = rs->inc_next;

    // descriptors of the member's first tiles.  When the blobs of all the
    // member's tiles (at most kTileCache) fit at the top of the tile area
    // with room below them for the largest work area (positions and force
    // vectors) of those tiles, they stay resident there, back to back.  With
    // two nodes per thread (fewer registers to spare) only a member's single
    // tile stays, where a loaded blob goes.
    const int n_mine = T.member < C.n_tiles ? (C.n_tiles - T.member + T.g - 1) / T.g : 0;
    int res_bytes = 0, res_work = 0;
    for (int k = 0, ti = T.member; k < kTileCache && ti < C.n_tiles; k++, ti += T.g) {
        const TileDev *tg = A.tiles + C.tile_base + ti;
        constexpr int kWords = (int)(offsetof(TileDev, lds_halo) / sizeof(int));
        if (threadIdx.x < kWords)
            reinterpret_cast<int *>(L.ctd + k)[threadIdx.x] =
                reinterpret_cast<const int *>(tg)[threadIdx.x];
        res_bytes += align16(tg->blob_bytes);
        res_work = max(res_work, tile_work_bytes(tg->n_own + tg->n_halo, tg->nb, tg->nt));
    }
    const bool resident = n_mine >= 1 && n_mine <= kTileCache &&
                          (NPT == 1 ? res_bytes + res_work <= A.tile_cap : n_mine == 1);
    // the resident blobs, or else the halo ids of the member's first tiles
    // (halo ids lead the blob), packed into the pool while they fit
    for (int k = 0, ti = T.member, used = 0, o = A.tile_cap; k < kTileCache && ti < C.n_tiles;
         k++, ti += T.g) {
        const TileDev *tg = A.tiles + C.tile_base + ti;
        if (resident) {
            o = NPT == 1 ? o - align16(tg->blob_bytes) : 0;
            if (threadIdx.x == 0) {
                L.ctd[k].lds_halo = -1;
                L.ctd[k].lds_blob = o;
            }
            const uint4 *__restrict__ src = A.blob + tg->blob16;
            uint4 *dst = reinterpret_cast<uint4 *>(L.blob + o);
            const int n16 = tg->blob_bytes >> 4;
            for (int i = threadIdx.x; i < n16; i += B) dst[i] = src[i];
            continue;
        }
        const int n_halo = tg->n_halo;
        const bool fits = used + n_halo <= kCachePool;
        if (threadIdx.x == 0) {
            L.ctd[k].lds_halo = fits ? used : -1;
            L.ctd[k].lds_blob = -1;
        }
        if (fits) {
            const int *src = reinterpret_cast<const int *>(A.blob + tg->blob16);
            for (int i = threadIdx.x; i < n_halo; i += B) L.chalo[used + i] = src[i];
            used += n_halo;
        }
    }
    __syncthreads();

    Fire fire;
    fire.nd = nd;
    StepScal S;
    S.dt = 0.0;
    S.dtsq = 0.0;
    S.dtm = 0.0;
    S.pend_x = 0.0;
    S.pend_y = 0.0;
    Red4 r = {0.0, 0.0, 0.0, 0.0};
    Red4 part;
    int status = NR_NOT_CONVERGED;
    bool paused = false;
    bool host_step_used = false;
    int step, steps32;
    double ff, fmax2;

    for (;;) {  // strain increments
        const bool fresh = phase == kRunFresh;
        if (fresh) {
            // strain step set by the host (first increment of this launch
            // only; the record is cleared when the job's visit ends, after
            // every member has read it) plus the next increment of the list
            double pend_x = 0.0, pend_y = 0.0;
            if (!host_step_used) {
                pend_x = nd->pend_x;
                pend_y = nd->pend_y;
                host_step_used = true;
            }
            if (inc_next < nd->inc_count) {
                const double2 d = A.incs[nd->inc_base + inc_next];
                pend_x += d.x;
                pend_y += d.y;
                inc_next++;
            }
            // the strain step advances the Lees-Edwards image shift
            C.shift = shift + (C.H > 0 ? pend_x : 0.0);
            S.pend_x = pend_x;
            S.pend_y = pend_y;
            S.alpha_old = 0.0;
            S.cmix = 0.0;
            S.mix = kMixNone;
            S.s = 0;
            if (lead) rs->t_start = wall_clock64();
            step = 0;
            steps32 = 0;
            fire.start();
            phase = CG ? kRunCg : A.mixed && max_steps > 0 ? kRun32 : kRun64;
        } else {
            C.shift = shift;
            step = rs->step;
            steps32 = rs->steps32;
            ff = rs->ff;
            fmax2 = rs->fmax2;
            fire.alpha = rs->alpha;
            fire.dt = rs->dt;
            fire.since = rs->since;
            S.alpha_old = rs->alpha_old;
            S.cmix = rs->cmix;
            S.mix = rs->mix;
            S.s = rs->s;
        }

        if (!CG && phase == kRun32) {
            // FP32 stage: anchor = positions after the strain step, in x[p ^ 1]
            const int pa = p ^ 1;
            double best, fswitch;
            int since_best;
            if (fresh) {
                part = pass32<B, NPT, kPass32Init>(A, T, C, S, p, pa, L);
                if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, r, s_red, s_flag)) status = NR_ABORTED;
                S.s = 1;
                S.pend_x = 0.0;
                S.pend_y = 0.0;
                fswitch = fmax(nd->fcut, A.switch_ratio * sqrt(r.m));
                best = r.m;
                since_best = 0;
            } else {
                fswitch = rs->fswitch;
                best = rs->best;
                since_best = rs->since_best;
            }
            while (status != NR_ABORTED && step < max_steps) {
                if (used >= budget) {
                    paused = true;
                    break;
                }
                step++;
                used++;
                fire.load(S);
                part = pass32<B, NPT, kPass32Step>(A, T, C, S, p, pa, L);
                if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, r, s_red, s_flag)) {
                    status = NR_ABORTED;
                    break;
                }
                S.s ^= 1;
                fire.update(r, S);
                if (!isfinite(r.c) || sqrt(r.m) < fswitch) break;
                if (r.m < best) {
                    best = r.m;
                    since_best = 0;
                } else if (++since_best > A.stall_steps) {
                    break;
                }
            }
            if (paused) {
                if (lead) {
                    rs->fswitch = fswitch;
                    rs->best = best;
                    rs->since_best = since_best;
                }
            } else if (status != NR_ABORTED) {
                // FP64 polish from anchor + displacement; FIRE continues with
                // the velocity carried over by the fuse
                steps32 = step;
                p = pa;
                part = pass<B, NPT, kPassFuse>(A, T, C, S, p, L);
                if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, r, s_red, s_flag)) status = NR_ABORTED;
                p ^= 1;
                ff = r.c;
                fmax2 = r.m;
                phase = kRun64;
                if (steps32 == 0) {
                    fire.start();
                    S.mix = kMixNone;
                }
            }
        } else if (fresh) {
            part = pass<B, NPT, kPassInit>(A, T, C, S, p, L);
            if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, r, s_red, s_flag)) status = NR_ABORTED;
            S.pend_x = 0.0;
            S.pend_y = 0.0;
            p ^= 1;
            ff = r.c;
            fmax2 = r.m;
        }
        shift = C.shift;

        if (CG && phase == kRunCg && status != NR_ABORTED) {
            // Nonlinear CG (Fletcher-Reeves or Polak-Ribiere+, secant line
            // search on phi'(a) = -F(x + a d).d): the algorithm of the FP64
            // oracle's CG.
            // f.d and |d|^2 of the new direction d = f + beta d_old follow from
            // the last accepted trial's reductions.
            const double fcut = nd->fcut;
            const int crit = nd->crit;
            const bool pr = nd->method == NR_METHOD_CG_PR;
            const double sqrt_dof = sqrt((double)(2 * C.n_nodes));
            double beta_next, step_len, dd_old, fdt_acc;
            int pd;
            if (fresh) {
                beta_next = 0.0;
                step_len = 1e-2;
                dd_old = 0.0;
                fdt_acc = 0.0;
                pd = p;  // zeroed by the init pass
            } else {
                beta_next = rs->alpha;
                step_len = rs->dt;
                dd_old = rs->alpha_old;
                fdt_acc = rs->cmix;
                pd = rs->s;
            }
            for (;;) {
                const bool conv =
                    crit == NR_CRIT_MAX ? sqrt(fmax2) < fcut : sqrt(ff) / sqrt_dof < fcut;
                if (conv || !isfinite(ff) || step >= max_steps) break;
                if (used >= budget) {
                    paused = true;
                    break;
                }
                step++;
                used++;
                double beta = beta_next;
                double fd = fma(beta, fdt_acc, ff);
                double dd = fma(beta * beta, dd_old, fma(2.0 * beta, fdt_acc, ff));
                if (!(fd > 0)) {  // not a descent direction: restart along f
                    beta = 0.0;
                    fd = ff;
                    dd = ff;
                }
                const double g0 = -fd;
                double a_prev = 0.0, g_prev = g0, a = step_len / sqrt(dd);
                Red4 tr = {0.0, 0.0, 0.0, 0.0};
                S.cg_beta = beta;
                S.pd = pd;
                int ls = 0;
                for (; ls < kCgLineSearch; ls++) {
                    S.cg_a = a;
                    S.cg_first = ls == 0;
                    part = pass<B, NPT, kPassCgTrial>(A, T, C, S, p, L);
                    if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, tr, s_red, s_flag)) {
                        status = NR_ABORTED;
                        break;
                    }
                    const double g = -tr.a;
                    if (fabs(g) <= kCgEta * fabs(g0)) break;
                    double an = a - g * (a - a_prev) / (g - g_prev);
                    if (!(an > 0) || !isfinite(an)) an = 2.0 * a;
                    if (an > 4.0 * a) an = 4.0 * a;
                    a_prev = a;
                    g_prev = g;
                    a = an;
                }
                if (status == NR_ABORTED) break;
                if (ls == kCgLineSearch) {  // evaluate the last step length
                    S.cg_a = a;
                    S.cg_first = 0;
                    part = pass<B, NPT, kPassCgTrial>(A, T, C, S, p, L);
                    if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, tr, s_red, s_flag)) {
                        status = NR_ABORTED;
                        break;
                    }
                }
                // accept the trial point; its direction becomes d_old
                step_len = a * sqrt(dd);
                p ^= 1;
                pd ^= 1;
                // Fletcher-Reeves |F|^2 / |F_old|^2, or Polak-Ribiere+
                // max(0, (|F|^2 - F.F_old) / |F_old|^2)
                beta_next = pr ? fmax(0.0, (tr.c - tr.b) / ff) : tr.c / ff;
                fdt_acc = tr.a;
                dd_old = dd;
                ff = tr.c;
                fmax2 = tr.m;
            }
            if (paused && lead) {
                rs->alpha = beta_next;
                rs->dt = step_len;
                rs->alpha_old = dd_old;
                rs->cmix = fdt_acc;
                rs->s = pd;
            }
        }

        if (!CG && phase == kRun64 && !paused && status != NR_ABORTED) {
            // run_fire_md :624-718
            const double fcut = nd->fcut;
            const int crit = nd->crit;
            const double sqrt_dof = sqrt((double)(2 * C.n_nodes));
            while (step < max_steps) {
                if (used >= budget) {
                    paused = true;
                    break;
                }
                step++;
                used++;
                fire.load(S);
                part = pass<B, NPT, kPassStep>(A, T, C, S, p, L);
                if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, r, s_red, s_flag)) {
                    status = NR_ABORTED;
                    break;
                }
                p ^= 1;
                ff = r.c;
                fmax2 = r.m;
                fire.update(r, S);
                const bool conv =
                    crit == NR_CRIT_MAX ? sqrt(fmax2) < fcut : sqrt(ff) / sqrt_dof < fcut;
                if (conv) break;
                if (!isfinite(ff)) break;
            }
        }
        if (paused) {
            status = NR_IN_PROGRESS;
            break;
        }
        const double sqrt_dof = sqrt((double)(2 * C.n_nodes));
        if (status != NR_ABORTED) {
            if (!isfinite(ff)) {
                status = NR_NONFINITE;
            } else {
                const double meas = nd->crit == NR_CRIT_MAX ? sqrt(fmax2) : sqrt(ff) / sqrt_dof;
                status = meas < nd->fcut ? NR_CONVERGED : NR_NOT_CONVERGED;
            }
        }
        // energies at the final positions (:723-757)
        Red4 e = {0.0, 0.0, 0.0, 0.0};
        if (status != NR_ABORTED) {
            part = pass<B, NPT, kPassEnergy>(A, T, C, S, p, L);
            if (!reduce4<B, poll_sweeps(NPT)>(A, T, part, e, s_red, s_flag)) status = NR_ABORTED;
        }
        if (lead) {
            NetOut o;
            o.rms = sqrt(ff) / sqrt_dof;
            o.fmax = sqrt(fmax2);
            o.es = C.ks > 0 ? e.a : 0.0;
            o.eb = C.kb > 0 ? e.b : 0.0;
            o.sxy = e.c / (C.W * C.yrange);
            o.steps = step;
            o.status = status;
            o.fp64_steps = step - steps32;
            o.pad = 0;
            o.t_start = rs->t_start;
            o.t_done = wall_clock64();
            A.out[net_id] = o;
            if (inc_next > 0 && inc_next <= nd->inc_count) A.inc_out[nd->out_base + inc_next - 1] = o;
        }
        phase = kRunFresh;
        // the next increment of the list starts from these positions
        if ((status == NR_CONVERGED || status == NR_NOT_CONVERGED) && inc_next < nd->inc_count)
            continue;
        break;
    }
    if (lead) {
        if (paused) {
            rs->step = step;
            rs->steps32 = steps32;
            rs->ff = ff;
            rs->fmax2 = fmax2;
            if (phase != kRunCg) {
                rs->alpha = fire.alpha;
                rs->dt = fire.dt;
                rs->since = fire.since;
                rs->alpha_old = S.alpha_old;
                rs->cmix = S.cmix;
                rs->mix = S.mix;
                rs->s = S.s;
            }
            NetOut o = {};
            o.rms = sqrt(ff) / sqrt((double)(2 * C.n_nodes));
            o.fmax = sqrt(fmax2);
            o.steps = step;
            o.status = NR_IN_PROGRESS;
            o.fp64_steps = step - steps32;
            o.t_start = rs->t_start;
            o.t_done = wall_clock64();
            A.out[net_id] = o;
        }
        rs->phase = phase;
        rs->inc_next = inc_next;
        nd->parity = p;
        nd->le_shift = shift;
        nd->pend_x = 0.0;
        nd->pend_y = 0.0;
    }
}

template <int B, int NPT, bool CG>
__device__ __forceinline__ void kernel_body(const KArgs &A)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double s_red[4 * (B / 64 + 1)];
    __shared__ int s_flag;
    __shared__ int s_job;
    Lds L;
    L.ctd = reinterpret_cast<TileDev *>(smem);
    L.chalo = reinterpret_cast<int *>(smem + kTileCache * sizeof(TileDev));
    L.blob = smem + kCacheBytes;
    L.pos = nullptr;  // placed per tile, after its blob
    L.fv = nullptr;

    const BlockRole role = A.roles[blockIdx.x];
    if (role.team < 0) return;  // idle workgroup
    Team T;
    T.member = role.member;
    T.g = role.g;
    T.part = role.part;
    T.rtag = 0;
    int heap = 0;  // position in the team's split tree (0: the whole team)
    T.id = role.tree;
    int cls = role.cls;
    // barrier generation of the (sub)team's words (they persist across launches)
    T.gen = 0;
    if (T.g > 1 && threadIdx.x == 0)
        T.gen = __hip_atomic_load(A.bar + (size_t)T.id * kBarStride + 16, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);

    for (;;) {
        int *queue = A.queues + 32 * cls;
        const int *jobs = A.class_jobs + A.class_off[cls];
        const int n_jobs = A.class_off[cls + 1] - A.class_off[cls];
        if (threadIdx.x == 0 && T.member == 0) {
            const int j = __hip_atomic_fetch_add(queue, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        if (job < n_jobs) {
            relax_network<B, NPT, CG>(A, T, jobs[job], L, s_red, &s_flag);
            if (__hip_atomic_load(A.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
            continue;
        }
        // This class has no job left (every member saw the same answer): the
        // team halves, each half serving the class of teams of its size, so
        // no workgroup idles while smaller classes still queue work.  A job
        // of a class is still relaxed by a team of that class's size with
        // members 0..g-1, so its arithmetic does not change.
        int next = -1;
        while (T.g > 1 && next < 0) {
            const int half = T.g >> 1, sub = T.member >= half ? 1 : 0;
            T.member -= sub * half;
            T.part += sub * half;
            T.g = half;
            heap = 2 * heap + 1 + sub;
            int lg = 0;
            while ((1 << lg) < half) lg++;
            next = lg < kMaxLog2Team ? A.cls_of_log2g[lg] : -1;
        }
        if (next < 0) return;
        cls = next;
        T.id = role.tree + heap;
        T.gen = 0;
        if (T.g > 1 && threadIdx.x == 0)
            T.gen = __hip_atomic_load(A.bar + (size_t)T.id * kBarStride + 16, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    }
}

// FIRE (and mixed precision) and CG entry points; CG is held to two waves
// per SIMD (256 VGPRs), the occupancy the FIRE kernel reaches by itself.
template <int B, int NPT>
__global__ __launch_bounds__(B) void relax_kernel(KArgs A)
{
    kernel_body<B, NPT, false>(A);
}

template <int B, int NPT>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(2))) void cg_kernel(KArgs A)
{
    kernel_body<B, NPT, true>(A);
}

// Affine steps of one strain step, applied on the device: network k's
// pending step grows by d[k] (graph replays repeat it without host writes).
__global__ void strain_step_kernel(NetDev *nets, int count, const double2 *d)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) {
        nets[k].pend_x += d[k].x;
        nets[k].pend_y += d[k].y;
    }
}

// Zero fill of n 8-byte words (a kernel, so that graphs hold no memset node).
__global__ void zero_words_kernel(unsigned long long *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

const void *kernel_for(int block, int npt, int method)
{
    const bool cg = method == NR_METHOD_CG || method == NR_METHOD_CG_PR;
    if (block == 256 && npt == 1)
        return cg ? (const void *)cg_kernel<256, 1> : (const void *)relax_kernel<256, 1>;
    if (block == 256 && npt == 2)
        return cg ? (const void *)cg_kernel<256, 2> : (const void *)relax_kernel<256, 2>;
    if (block == 512 && npt == 1)
        return cg ? (const void *)cg_kernel<512, 1> : (const void *)relax_kernel<512, 1>;
    if (block == 512 && npt == 2)
        return cg ? (const void *)cg_kernel<512, 2> : (const void *)relax_kernel<512, 2>;
    return nullptr;
}

}  // namespace

int launch_relax(const KArgs &a, int block, int npt, int method, int grid, size_t lds,
                 bool cooperative, void *stream)
{
    const void *fn = kernel_for(block, npt, method);
    if (!fn) return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    void *args[] = {const_cast<KArgs *>(&a)};
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
    if (cooperative)
        e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args, (unsigned)lds, s);
    else
        e = hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, s);
    return (int)e;
}

int max_resident_blocks(int block, int npt, size_t lds, int *per_cu, int *n_cu)
{
    // the smaller residency of the FIRE and CG kernels
    int nb = 1 << 30;
    for (int method = NR_METHOD_FIRE; method <= NR_METHOD_CG; method++) {
        const void *f = kernel_for(block, npt, method);
        if (!f) return (int)hipErrorInvalidValue;
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return (int)e;
        int n = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, lds);
        if (e != hipSuccess) return (int)e;
        nb = std::min(nb, n);
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return (int)e;
    *per_cu = nb;
    *n_cu = prop.multiProcessorCount;
    return 0;
}

int launch_strain_step(NetDev *nets, int count, const double2 *d, void *stream)
{
    if (count <= 0) return 0;
    hipLaunchKernelGGL(strain_step_kernel, dim3((count + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, nets, count, d);
    return (int)hipGetLastError();
}

int launch_zero(void *p, size_t bytes, void *stream)
{
    const size_t n = bytes / 8;
    if (n == 0) return 0;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(zero_words_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<unsigned long long *>(p), n);
    return (int)hipGetLastError();
}

}  // namespace nr
