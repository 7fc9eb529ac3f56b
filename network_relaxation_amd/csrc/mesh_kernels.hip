// gfx950 kernels of the network + FEM-gel composite relax (the reference's
// network_mesh_relax.cu run_fire_md :865-1066): FIRE in FP64 on one
// composite, every vector and scalar resident on the device.  One FIRE step
// is two launches:
//   mesh_tile      the elements of spatial tiles into LDS, the gather, the
//                  second half kick and the partials (below; the element
//                  forces never go through HBM)
//   mesh_fire      as below
// or, unfused (NRELAX_MESH_TILES=0 at set-up), three:
//   mesh_elements  one thread per bond, triple and facet; each writes its
//                  forces into its own slots (one array per bond, triple end
//                  and facet corner, in element order: contiguous writes, no
//                  atomics)
//   mesh_gather    one thread per point: the clamped sum of the slots that
//                  name it (its twin's too when the gel couples twins: the
//                  reference's duplicate_forces :754-779 as a gather), the
//                  second half kick, partials of v.F, v.v, F.F, max |F|^2
//                  (the last workgroup to finish sums the partials in one
//                  fixed order)
//   mesh_fire      mixes v, adapts dt and alpha, tests convergence and
//                  drifts x and v into the next step
// Scalars live in two parity slots (a step reads slot par, its mesh_fire
// writes slot par ^ 1), so a fixed sequence of steps can be captured once and
// replayed; once a slot's done flag is up every kernel returns at once and
// mesh_fire carries the flag forward.
#include <hip/hip_runtime.h>

#include "mesh_device.hpp"


namespace nr {
namespace {

constexpr int kThreads = 256;
constexpr int kXcds = 8;  // MI355X: workgroup b is dispatched to XCD b % 8

// Chunk of workgroup b among g: each XCD takes a contiguous range of chunks,
// so neighbouring points (which share element slots) meet in one XCD's L2.
// A bijection on [0, g) for any g.
__device__ inline int xcd_chunk(int b, int g)
{
    const int x = b % kXcds, q = b / kXcds, per = g / kXcds, rem = g % kXcds;
    return x * per + min(x, rem) + q;
}

__device__ inline double2 ld2(const double *x, int i)
{
    return *reinterpret_cast<const double2 *>(x + 2 * i);
}

__device__ inline void st2(double *x, int i, double a, double b)
{
    *reinterpret_cast<double2 *>(x + 2 * i) = make_double2(a, b);
}

__global__ __launch_bounds__(kThreads) void mesh_elements(MeshDev m, int par)
{
    if (m.st[par * kStSlot + kStDone] != 0.0) return;
    const int total = m.nb + m.nt + m.nf;
    const int chunk = xcd_chunk(blockIdx.x, gridDim.x);
    for (int k = chunk * kThreads + threadIdx.x; k < total; k += gridDim.x * kThreads) {
        if (k < m.nb) {  // axial_forces :534-562
            const int2 ij = m.bond[k];
            const double2 ro = m.bond_geo[k];  // (rest length, x image offset)
            const double2 pi = ld2(m.x, ij.x), pj = ld2(m.x, ij.y);
            const double rx = pj.x + ro.y - pi.x, ry = pj.y - pi.y;
            const double d2 = fma(rx, rx, ry * ry);
            const double s = m.ks * (1.0 - ro.x * (1.0 / sqrt(d2)));
            m.slot[k] = make_double2(s * rx, s * ry);
        } else if (k < m.nb + m.nt) {  // bending_forces :565-617
            const int t = k - m.nb;
            const int4 hab = m.tri[t];
            const double2 oo = m.tri_off[t];  // x image offsets of the two arms
            const double2 ph = ld2(m.x, hab.x), pa = ld2(m.x, hab.y), pb = ld2(m.x, hab.z);
            const double ax = pa.x + oo.x - ph.x, ay = pa.y - ph.y;
            const double bx = pb.x + oo.y - ph.x, by = pb.y - ph.y;
            const double ia = 1.0 / fma(ax, ax, ay * ay), ib = 1.0 / fma(bx, bx, by * by);
            const double cr = fma(ax, by, -(ay * bx)), dt = fma(ax, bx, ay * by);
            const double g = m.kb * atan2(-cr, -dt);  // kb (theta - pi)
            const double f2 = g * ia, f3 = g * ib;
            m.slot[m.nb + t] = make_double2(f2 * -ay, f2 * ax);
            m.slot[m.nb + m.nt + t] = make_double2(f3 * by, -f3 * bx);
        } else {  // gel_forces :702-750
            const int f = k - m.nb - m.nt;
            const int4 v = m.facet[f];
            const double *M = m.inv + 6 * (size_t)f;
            const double A = m.area[f];
            const double2 p1 = ld2(m.x, v.x), p2 = ld2(m.x, v.y), p3 = ld2(m.x, v.z);
            const double m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3], m4 = M[4], m5 = M[5];
            const double f11 = fma(p3.x, m4, fma(p2.x, m2, p1.x * m0));
            const double f12 = fma(p3.x, m5, fma(p2.x, m3, p1.x * m1));
            const double f21 = fma(p3.y, m4, fma(p2.y, m2, p1.y * m0));
            const double f22 = fma(p3.y, m5, fma(p2.y, m3, p1.y * m1));
            const double J = fma(f11, f22, -(f12 * f21));
            const double I1 = fma(f11, f11, fma(f12, f12, fma(f21, f21, f22 * f22)));
            const double c1 = -m.smod * A / J;
            const double c2 = (m.bmod * (m.pref_d + 1.0 - J) + m.smod * I1 / (2.0 * J * J)) * A;
            const double d11 = fma(f11, c1, f22 * c2), d12 = fma(f12, c1, -(f21 * c2));
            const double d21 = fma(f21, c1, -(f12 * c2)), d22 = fma(f22, c1, f11 * c2);
            double2 *o = m.slot + m.nb + 2 * (size_t)m.nt + f;  // one array per corner
            o[0] = make_double2(fma(d12, m1, d11 * m0), fma(d22, m1, d21 * m0));
            o[m.nf] = make_double2(fma(d12, m3, d11 * m2), fma(d22, m3, d21 * m2));
            o[2 * (size_t)m.nf] = make_double2(fma(d12, m5, d11 * m4), fma(d22, m5, d21 * m4));
        }
    }
}

__device__ inline void add_entry(double &sx, double &sy, int e, double2 s)
{
    sx += (e & 1) ? -s.x : s.x;
    sy += (e & 1) ? -s.y : s.y;
}

// Sum of the slots that name point i, in entry order (bonds, then triples,
// then facets); entries are (slot << 1) | negate.  Each section's entries
// are loaded a batch at a time before their slots, and the slot loads before
// the first sum, so a point costs a few memory round trips, not two per
// entry; column-major sections make each load coalesced across the wave.
constexpr int kEntBatch = 8;

__device__ inline void sum_section(const MeshDev &m, long base, int count, int i, double &sx,
                                   double &sy)
{
    for (int k = 0; k < count; k += kEntBatch) {
        int e[kEntBatch];
        double2 v[kEntBatch];
#pragma unroll
        for (int j = 0; j < kEntBatch; j++)
            e[j] = k + j < count ? m.ent[base + (long)(k + j) * m.n + i] : -1;
#pragma unroll
        for (int j = 0; j < kEntBatch; j++)
            v[j] = e[j] >= 0 ? m.slot[e[j] >> 1] : make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < kEntBatch; j++)
            if (e[j] >= 0) add_entry(sx, sy, e[j], v[j]);
    }
}

__device__ inline double2 point_sum(const MeshDev &m, int i)
{
    double sx = 0.0, sy = 0.0;
    const int c = m.cnt[i];
    sum_section(m, m.sec[0], c & 255, i, sx, sy);
    sum_section(m, m.sec[1], (c >> 8) & 255, i, sx, sy);
    sum_section(m, m.sec[2], (c >> 16) & 255, i, sx, sy);
    return make_double2(sx, sy);
}

template <int B>
__device__ inline double block_sum(double v, double *s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = B / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

template <int B>
__device__ inline double block_max(double v, double *s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = B / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

// Sums (a, b, c) and max (d) over the workgroup: wave butterflies, then the
// waves in order; every thread gets the results.  Two barriers.
__device__ inline void block_reduce4(double &a, double &b, double &c, double &d, double *s)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
        d = fmax(d, __shfl_xor(d, o, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s[4 * wave] = a;
        s[4 * wave + 1] = b;
        s[4 * wave + 2] = c;
        s[4 * wave + 3] = d;
    }
    __syncthreads();
    a = b = c = d = 0.0;
    for (int w = 0; w < kThreads / 64; w++) {
        a += s[4 * w];
        b += s[4 * w + 1];
        c += s[4 * w + 2];
        d = fmax(d, s[4 * w + 3]);
    }
    __syncthreads();
}

// A workgroup's four partials go out as agent-scope stores (coherent across
// the XCDs without an L2 write-back); once they have landed the workgroup
// takes a ticket, so the workgroup that draws the last ticket reads every
// partial.  A device-wide fence per workgroup instead writes back the XCD's
// whole L2 (every dirty F and v line) each time.
__device__ inline unsigned publish_partials(const MeshDev &m, double P, double vv, double ff,
                                            double fm)
{
    unsigned long long *o = reinterpret_cast<unsigned long long *>(m.part + 4 * (size_t)blockIdx.x);
    const double w[4] = {P, vv, ff, fm};
#pragma unroll
    for (int k = 0; k < 4; k++)
        __hip_atomic_store(o + k, (unsigned long long)__double_as_longlong(w[k]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return __hip_atomic_fetch_add(m.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline double load_partial(const MeshDev &m, size_t i)
{
    return __longlong_as_double((long long)__hip_atomic_load(
        reinterpret_cast<unsigned long long *>(m.part) + i, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT));
}

// kick != 0: v += dt/(2m) F after the force (a step); 0: the initial force.
__global__ __launch_bounds__(kThreads) void mesh_gather(MeshDev m, int par, int kick)
{
    __shared__ double s[kThreads];
    const double *st = m.st + par * kStSlot;
    if (st[kStDone] != 0.0) return;
    const double dtm = 0.5 * st[kStDt] / m.mass;
    double P = 0.0, vv = 0.0, ff = 0.0, fm = 0.0;
    const int chunk = xcd_chunk(blockIdx.x, gridDim.x);
    for (int i = chunk * kThreads + threadIdx.x; i < m.n; i += gridDim.x * kThreads) {
        const unsigned char c = m.fcode[i];
        const double2 si = point_sum(m, i);
        double fx = (c & 1) ? si.x : 0.0, fy = (c & 2) ? si.y : 0.0;
        const int p = m.twin ? m.twin[i] : -1;
        if (p >= 0) {
            const unsigned char cp = m.fcode[p];
            const double2 sp = point_sum(m, p);
            // both twins get F_a + F_b (one value: the sum commutes)
            fx += (cp & 1) ? sp.x : 0.0;
            fy += (cp & 2) ? sp.y : 0.0;
        }
        st2(m.f, i, fx, fy);
        double2 v = ld2(m.v, i);
        if (kick) {
            v.x = fma(dtm, fx, v.x);
            v.y = fma(dtm, fy, v.y);
            st2(m.v, i, v.x, v.y);
        }
        P = fma(v.x, fx, fma(v.y, fy, P));
        vv = fma(v.x, v.x, fma(v.y, v.y, vv));
        const double q = fma(fx, fx, fy * fy);
        ff += q;
        fm = fmax(fm, q);
    }
    P = block_sum<kThreads>(P, s);
    vv = block_sum<kThreads>(vv, s);
    ff = block_sum<kThreads>(ff, s);
    fm = block_max<kThreads>(fm, s);
    __shared__ unsigned ticket;
    if (threadIdx.x == 0) ticket = publish_partials(m, P, vv, ff, fm);
    __syncthreads();
    if (ticket != gridDim.x - 1) return;
    // the last workgroup: every partial has landed; sum them in workgroup order
    P = vv = ff = fm = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads) {
        P += load_partial(m, 4 * (size_t)b);
        vv += load_partial(m, 4 * (size_t)b + 1);
        ff += load_partial(m, 4 * (size_t)b + 2);
        fm = fmax(fm, load_partial(m, 4 * (size_t)b + 3));
    }
    P = block_sum<kThreads>(P, s);
    vv = block_sum<kThreads>(vv, s);
    ff = block_sum<kThreads>(ff, s);
    fm = block_max<kThreads>(fm, s);
    if (threadIdx.x == 0) {
        m.tot[0] = P;
        m.tot[1] = vv;
        m.tot[2] = ff;
        m.tot[3] = fm;
        *m.ticket = 0u;  // ready for the next launch
    }
}

// Fused element + gather step over the tiles of MeshTile, persistent: the
// launch has at most as many workgroups as are resident, and workgroup b
// takes tiles b, b + G, ...  Per tile: the element records and the positions
// of the tile's points are loaded together (the descriptor, the blob rows
// and the point ids were fetched during the previous tile); the blob and
// the positions go into LDS; each element of the tile writes its force
// vectors into LDS (elements on a tile boundary are computed by every tile
// that needs them); each own point sums its entries in element order (its
// own, clamped; then its twin's, clamped by the twin's code), writes F,
// takes the second half kick (kick != 0) and adds to the workgroup's
// partials of v.F, v.v, F.F and max |F|^2.  The last workgroup to finish
// sums the partials in workgroup order.  Element forces are those of
// mesh_elements, the entry order that of mesh_gather, so both paths give the
// same forces bit for bit.
struct TilePre {
    MeshTile d;
    uint4 rows[kTileBlobRows];
    int lid[kTileLocalPerThread];
};

__device__ inline void tile_prefetch(const MeshDev &m, int t, TilePre &p)
{
    p.d = m.tiles[t];
    const MeshBlob L = mesh_blob(p.d.n_own, p.d.n_halo, p.d.n_inc);
    const uint4 *src = m.tblob + p.d.blob16;
    const int n16 = L.bytes >> 4, n_loc = p.d.n_own + p.d.n_halo;
#pragma unroll
    for (int j = 0; j < kTileBlobRows; j++) {
        const int i = threadIdx.x + j * kThreads;
        if (i < n16) p.rows[j] = src[i];
    }
    const int *gids = reinterpret_cast<const int *>(src) + L.ids / 4;
#pragma unroll
    for (int j = 0; j < kTileLocalPerThread; j++) {
        const int k = threadIdx.x + j * kThreads;
        p.lid[j] = k < n_loc ? gids[k] : 0;
    }
}

__global__ __launch_bounds__(kThreads) void mesh_tile(MeshDev m, int par, int kick)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double s[kThreads];
    const double *st = m.st + par * kStSlot;
    if (st[kStDone] != 0.0) return;
    const double dtm = 0.5 * st[kStDt] / m.mass;
    const int tid = threadIdx.x;
    double P = 0.0, vv = 0.0, ff = 0.0, fm = 0.0;
    TilePre cur;
    int t = blockIdx.x;
    if (t < m.n_tiles) tile_prefetch(m, t, cur);
    for (; t < m.n_tiles; t += gridDim.x) {
        const MeshTile td = cur.d;
        const MeshBlob L = mesh_blob(td.n_own, td.n_halo, td.n_inc);
        const int n_loc = td.n_own + td.n_halo;
        // element records of this thread (at most kTileBondsPerThread bonds
        // and so on, a bound the planner keeps), own velocity and positions
        unsigned bij[kTileBondsPerThread];
        double2 bgeo[kTileBondsPerThread];
#pragma unroll
        for (int j = 0; j < kTileBondsPerThread; j++) {
            const int e = tid + j * kThreads;
            if (e < td.nb) {
                bij[j] = m.tb_ij[td.b0 + e];
                bgeo[j] = m.tb_geo[td.b0 + e];
            }
        }
        uint2 tidx[kTileTriplesPerThread];
        double2 toff[kTileTriplesPerThread];
#pragma unroll
        for (int j = 0; j < kTileTriplesPerThread; j++) {
            const int e = tid + j * kThreads;
            if (e < td.nt) {
                tidx[j] = m.tt_idx[td.t0 + e];
                toff[j] = m.tt_off[td.t0 + e];
            }
        }
        uint2 fidx[kTileFacetsPerThread];
        double2 fm0[kTileFacetsPerThread], fm1[kTileFacetsPerThread], fm2[kTileFacetsPerThread];
        double farea[kTileFacetsPerThread];
#pragma unroll
        for (int j = 0; j < kTileFacetsPerThread; j++) {
            const int e = tid + j * kThreads;
            if (e < td.nf) {
                const size_t g = (size_t)(td.f0 + e);
                fidx[j] = m.tf_idx[g];
                fm0[j] = m.tf_m[3 * g];
                fm1[j] = m.tf_m[3 * g + 1];
                fm2[j] = m.tf_m[3 * g + 2];
                farea[j] = m.tf_area[g];
            }
        }
        const bool own = tid < td.n_own;  // at most kThreads own points
        const int gid = own ? cur.lid[0] : 0;
        double2 v = make_double2(0.0, 0.0);
        if (own) v = ld2(m.v, gid);
        double2 lx[kTileLocalPerThread];
#pragma unroll
        for (int j = 0; j < kTileLocalPerThread; j++)
            if (tid + j * kThreads < n_loc) lx[j] = ld2(m.x, cur.lid[j]);
        // LDS is free (the previous tile ended with a barrier)
        {
            uint4 *dst = reinterpret_cast<uint4 *>(smem);
            const int n16 = L.bytes >> 4;
#pragma unroll
            for (int j = 0; j < kTileBlobRows; j++) {
                const int i = tid + j * kThreads;
                if (i < n16) dst[i] = cur.rows[j];
            }
        }
        double2 *pos = reinterpret_cast<double2 *>(smem + L.bytes);
        double2 *fv = pos + n_loc;
#pragma unroll
        for (int j = 0; j < kTileLocalPerThread; j++)
            if (tid + j * kThreads < n_loc) pos[tid + j * kThreads] = lx[j];
        // the next tile's descriptor, blob rows and point ids travel while
        // this one computes
        if (t + (int)gridDim.x < m.n_tiles) tile_prefetch(m, t + gridDim.x, cur);
        __syncthreads();
        const uint16_t *ioff = reinterpret_cast<const uint16_t *>(smem + L.ioff);
        const uint16_t *imid = reinterpret_cast<const uint16_t *>(smem + L.imid);
        const uint8_t *mask = smem + L.mask;
        const uint16_t *inc = reinterpret_cast<const uint16_t *>(smem + L.inc);
        // bonds (axial_forces :534-562)
#pragma unroll
        for (int j = 0; j < kTileBondsPerThread; j++) {
            const int e = tid + j * kThreads;
            if (e < td.nb) {
                const unsigned ij = bij[j];
                const double2 ro = bgeo[j];  // (rest length, x image offset)
                const double2 pi = pos[ij & 0xFFFF], pj = pos[ij >> 16];
                const double rx = pj.x + ro.y - pi.x, ry = pj.y - pi.y;
                const double d2 = fma(rx, rx, ry * ry);
                const double sc = m.ks * (1.0 - ro.x * (1.0 / sqrt(d2)));
                fv[e] = make_double2(sc * rx, sc * ry);
            }
        }
        // triples (bending_forces :565-617)
#pragma unroll
        for (int j = 0; j < kTileTriplesPerThread; j++) {
            const int e = tid + j * kThreads;
            if (e < td.nt) {
                const uint2 hab = tidx[j];
                const double2 oo = toff[j];  // x image offsets of the two arms
                const double2 ph = pos[hab.x & 0xFFFF], pa = pos[hab.x >> 16], pb = pos[hab.y];
                const double ax = pa.x + oo.x - ph.x, ay = pa.y - ph.y;
                const double bx = pb.x + oo.y - ph.x, by = pb.y - ph.y;
                const double ia = 1.0 / fma(ax, ax, ay * ay), ib = 1.0 / fma(bx, bx, by * by);
                const double cr = fma(ax, by, -(ay * bx)), dt = fma(ax, bx, ay * by);
                const double g = m.kb * atan2(-cr, -dt);  // kb (theta - pi)
                const double f2 = g * ia, f3 = g * ib;
                fv[td.nb + e] = make_double2(f2 * -ay, f2 * ax);
                fv[td.nb + td.nt + e] = make_double2(f3 * by, -f3 * bx);
            }
        }
        // facets (gel_forces :702-750)
#pragma unroll
        for (int j = 0; j < kTileFacetsPerThread; j++) {
            const int f = tid + j * kThreads;
            if (f < td.nf) {
                const uint2 c3 = fidx[j];
                const double A = farea[j];
                const double2 p1 = pos[c3.x & 0xFFFF], p2 = pos[c3.x >> 16], p3 = pos[c3.y];
                const double m0 = fm0[j].x, m1 = fm0[j].y, m2 = fm1[j].x, m3 = fm1[j].y,
                             m4 = fm2[j].x, m5 = fm2[j].y;
                const double f11 = fma(p3.x, m4, fma(p2.x, m2, p1.x * m0));
                const double f12 = fma(p3.x, m5, fma(p2.x, m3, p1.x * m1));
                const double f21 = fma(p3.y, m4, fma(p2.y, m2, p1.y * m0));
                const double f22 = fma(p3.y, m5, fma(p2.y, m3, p1.y * m1));
                const double J = fma(f11, f22, -(f12 * f21));
                const double I1 = fma(f11, f11, fma(f12, f12, fma(f21, f21, f22 * f22)));
                const double c1 = -m.smod * A / J;
                const double c2 = (m.bmod * (m.pref_d + 1.0 - J) + m.smod * I1 / (2.0 * J * J)) * A;
                const double d11 = fma(f11, c1, f22 * c2), d12 = fma(f12, c1, -(f21 * c2));
                const double d21 = fma(f21, c1, -(f12 * c2)), d22 = fma(f22, c1, f11 * c2);
                double2 *o = fv + td.nb + 2 * td.nt + f;  // one array per corner
                o[0] = make_double2(fma(d12, m1, d11 * m0), fma(d22, m1, d21 * m0));
                o[td.nf] = make_double2(fma(d12, m3, d11 * m2), fma(d22, m3, d21 * m2));
                o[2 * td.nf] = make_double2(fma(d12, m5, d11 * m4), fma(d22, m5, d21 * m4));
            }
        }
        __syncthreads();
        if (own) {
            const int lo = ioff[tid], mid = imid[tid], hi = ioff[tid + 1];
            double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
            for (int q = lo; q < mid; q++) {
                const int e = inc[q];
                add_entry(sx, sy, e, fv[e >> 1]);
            }
            for (int q = mid; q < hi; q++) {
                const int e = inc[q];
                add_entry(tx, ty, e, fv[e >> 1]);
            }
            const unsigned c = mask[tid];
            double fx = (c & 1) ? sx : 0.0, fy = (c & 2) ? sy : 0.0;
            if (c & 16) {  // both twins get F_a + F_b
                fx += (c & 4) ? tx : 0.0;
                fy += (c & 8) ? ty : 0.0;
            }
            st2(m.f, gid, fx, fy);
            if (kick) {
                v.x = fma(dtm, fx, v.x);
                v.y = fma(dtm, fy, v.y);
                st2(m.v, gid, v.x, v.y);
            }
            P = fma(v.x, fx, fma(v.y, fy, P));
            vv = fma(v.x, v.x, fma(v.y, v.y, vv));
            const double q2 = fma(fx, fx, fy * fy);
            ff += q2;
            fm = fmax(fm, q2);
        }
        __syncthreads();  // the next tile overwrites the LDS
    }
    block_reduce4(P, vv, ff, fm, s);
    __shared__ unsigned ticket;
    if (tid == 0) ticket = publish_partials(m, P, vv, ff, fm);
    __syncthreads();
    if (ticket != gridDim.x - 1) return;
    // the last workgroup: every partial has landed; sum them in workgroup order
    P = vv = ff = fm = 0.0;
    for (int b = tid; b < (int)gridDim.x; b += kThreads) {
        P += load_partial(m, 4 * (size_t)b);
        vv += load_partial(m, 4 * (size_t)b + 1);
        ff += load_partial(m, 4 * (size_t)b + 2);
        fm = fmax(fm, load_partial(m, 4 * (size_t)b + 3));
    }
    block_reduce4(P, vv, ff, fm, s);
    if (tid == 0) {
        m.tot[0] = P;
        m.tot[1] = vv;
        m.tot[2] = ff;
        m.tot[3] = fm;
        *m.ticket = 0u;  // ready for the next launch
    }
}

// start != 0: the drift of step 1 after the initial force (no mix, no test).
__global__ __launch_bounds__(kThreads) void mesh_fire(MeshDev m, int par, int start)
{
    const double *st = m.st + par * kStSlot;
    double *nx = m.st + (par ^ 1) * kStSlot;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (st[kStDone] != 0.0) {
        if (lead) nx[kStDone] = 1.0;
        return;
    }
    const double P = m.tot[0], vv = m.tot[1], ff = m.tot[2], fm = m.tot[3];

    double dt = st[kStDt], alpha = st[kStAlpha], since = st[kStSince];
    const double step = st[kStStep];
    const double fmag = sqrt(ff);
    bool mix = false, zero = false, done = false;
    double c = 0.0, a_mix = alpha;
    int status = kMeshRunning;
    if (!start) {
        c = alpha * sqrt(vv) / fmag;
        mix = true;
        if (P > 0) {
            since += 1.0;
            if (since > (double)m.nmin) {
                dt = fmin(dt * m.f_inc, m.dt_max);
                alpha *= m.f_alpha;
            }
        } else {
            since = 0.0;
            dt *= m.f_dec;
            alpha = m.alpha0;
            zero = true;
        }
        const double measure = m.criterion ? sqrt(fm) : fmag / m.sqrt_dof;
        if (measure < m.fcut) {
            done = true;
            status = kMeshConverged;
        } else if (!isfinite(ff)) {
            done = true;
            status = kMeshNonfinite;
        }
    }
    if (!done && step >= (double)m.max_steps) {
        done = true;
        status = kMeshNotConverged;
    }
    const double dtm = 0.5 * dt / m.mass, dtsq = 0.5 * dt * dt / m.mass;
    const int chunk = xcd_chunk(blockIdx.x, gridDim.x);
    for (int i = chunk * kThreads + threadIdx.x; i < m.n; i += gridDim.x * kThreads) {
        double2 v = ld2(m.v, i);
        const double2 f = ld2(m.f, i);
        if (zero) {
            v = make_double2(0.0, 0.0);
        } else if (mix) {
            v.x = fma(c, f.x, fma(-a_mix, v.x, v.x));
            v.y = fma(c, f.y, fma(-a_mix, v.y, v.y));
        }
        if (!done) {  // the next step's drift and first half kick
            double2 x = ld2(m.x, i);
            x.x = fma(dtsq, f.x, fma(dt, v.x, x.x));
            x.y = fma(dtsq, f.y, fma(dt, v.y, x.y));
            st2(m.x, i, x.x, x.y);
            v.x = fma(dtm, f.x, v.x);
            v.y = fma(dtm, f.y, v.y);
        }
        st2(m.v, i, v.x, v.y);
    }
    if (lead) {
        nx[kStDt] = dt;
        nx[kStAlpha] = alpha;
        nx[kStSince] = since;
        nx[kStStep] = done ? step : step + 1.0;
        nx[kStDone] = done ? 1.0 : 0.0;
        if (done) {
            double *r = m.st + 2 * kStSlot;
            r[kResRms] = fmag / m.sqrt_dof;
            r[kResFmax] = sqrt(fm);
            r[kResSteps] = step;
            r[kResStatus] = (double)status;
        }
    }
}

// Affine step of affine_displace :783-793 over the n points.
__global__ __launch_bounds__(kThreads) void mesh_affine(double *x, int n, double miny,
                                                        double yrange, double xdisp,
                                                        double ydisp)
{
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        double2 p = ld2(x, i);
        p.x += xdisp * (p.y - miny) / yrange;
        p.y += ydisp * (p.y - miny) / yrange;
        st2(x, i, p.x, p.y);
    }
}

// Energies of run_fire_md :1030-1056: per-workgroup partials of E_stretch,
// E_bend and E_gel (axial_energies :620-638, bending_energies :641-671,
// gel_energy :674-699).
__global__ __launch_bounds__(kThreads) void mesh_energy(MeshDev m, double *part)
{
    __shared__ double s[kThreads];
    double es = 0.0, eb = 0.0, eg = 0.0;
    const int total = m.nb + m.nt + m.nf;
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < total; k += gridDim.x * kThreads) {
        if (k < m.nb) {
            const int2 ij = m.bond[k];
            const double2 ro = m.bond_geo[k];
            const double2 pi = ld2(m.x, ij.x), pj = ld2(m.x, ij.y);
            const double rx = pj.x + ro.y - pi.x, ry = pj.y - pi.y;
            const double dl = sqrt(fma(rx, rx, ry * ry)) - ro.x;
            es += 0.5 * m.ks * dl * dl;
        } else if (k < m.nb + m.nt) {
            const int t = k - m.nb;
            const int4 hab = m.tri[t];
            const double2 oo = m.tri_off[t];
            const double2 ph = ld2(m.x, hab.x), pa = ld2(m.x, hab.y), pb = ld2(m.x, hab.z);
            const double ax = pa.x + oo.x - ph.x, ay = pa.y - ph.y;
            const double bx = pb.x + oo.y - ph.x, by = pb.y - ph.y;
            const double th = atan2(-fma(ax, by, -(ay * bx)), -fma(ax, bx, ay * by));
            eb += 0.5 * m.kb * th * th;
        } else {
            const int f = k - m.nb - m.nt;
            const int4 v = m.facet[f];
            const double *M = m.inv + 6 * (size_t)f;
            const double A = m.area[f];
            const double2 p1 = ld2(m.x, v.x), p2 = ld2(m.x, v.y), p3 = ld2(m.x, v.z);
            const double f11 = fma(p3.x, M[4], fma(p2.x, M[2], p1.x * M[0]));
            const double f12 = fma(p3.x, M[5], fma(p2.x, M[3], p1.x * M[1]));
            const double f21 = fma(p3.y, M[4], fma(p2.y, M[2], p1.y * M[0]));
            const double f22 = fma(p3.y, M[5], fma(p2.y, M[3], p1.y * M[1]));
            const double J = fma(f11, f22, -(f12 * f21));
            const double I1 = fma(f11, f11, fma(f12, f12, fma(f21, f21, f22 * f22)));
            const double dJ = J - 1.0 - m.pref_d;
            eg += m.smod * A * (I1 / J - 2.0) / 2.0 + m.bmod * A * dJ * dJ / 2.0;
        }
    }
    es = block_sum<kThreads>(es, s);
    eb = block_sum<kThreads>(eb, s);
    eg = block_sum<kThreads>(eg, s);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = es;
        part[3 * blockIdx.x + 1] = eb;
        part[3 * blockIdx.x + 2] = eg;
    }
}

}  // namespace

int mesh_element_blocks(const MeshDev &m)
{
    const long total = (long)m.nb + m.nt + m.nf;
    return (int)std::max(1L, std::min((long)kMeshMaxBlocks, (total + kThreads - 1) / kThreads));
}

int mesh_point_blocks(int n)
{
    return std::max(1, std::min(kMeshMaxBlocks, (n + kThreads - 1) / kThreads));
}

int launch_mesh_step(const MeshDev &m, int par, int phase, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    // phase 0: a whole step; 1: the initial force and the first drift
    if (m.n_tiles > 0) {
        // more than the default 64 KB of dynamic LDS needs the attribute
        const hipError_t e = hipFuncSetAttribute((const void *)mesh_tile,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 m.tile_lds);
        if (e != hipSuccess) return (int)e;
        hipLaunchKernelGGL(mesh_tile, dim3(m.tile_grid), dim3(kThreads), (size_t)m.tile_lds, st,
                           m, par, phase == 0 ? 1 : 0);
    } else {
        hipLaunchKernelGGL(mesh_elements, dim3(mesh_element_blocks(m)), dim3(kThreads), 0, st, m,
                           par);
        hipLaunchKernelGGL(mesh_gather, dim3(m.gather_blocks), dim3(kThreads), 0, st, m, par,
                           phase == 0 ? 1 : 0);
    }
    hipLaunchKernelGGL(mesh_fire, dim3(m.gather_blocks), dim3(kThreads), 0, st, m, par,
                       phase == 0 ? 0 : 1);
    return (int)hipGetLastError();
}

int launch_mesh_affine(double *x, int n, double miny, double yrange, double xdisp, double ydisp,
                       void *stream)
{
    hipLaunchKernelGGL(mesh_affine, dim3(mesh_point_blocks(n)), dim3(kThreads), 0,
                       (hipStream_t)stream, x, n, miny, yrange, xdisp, ydisp);
    return (int)hipGetLastError();
}

int launch_mesh_energy(const MeshDev &m, double *part, int blocks, void *stream)
{
    hipLaunchKernelGGL(mesh_energy, dim3(blocks), dim3(kThreads), 0, (hipStream_t)stream, m,
                       part);
    return (int)hipGetLastError();
}

}  // namespace nr
