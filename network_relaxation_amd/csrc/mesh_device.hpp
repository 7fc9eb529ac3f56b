// Device layout and launches of the network + FEM-gel composite relax
// (mesh_kernels.hip, composite.cpp).
#pragma once

#include <hip/hip_runtime.h>

namespace nr {

// Scalar state: two parity slots of kStSlot doubles, then the result.
constexpr int kStSlot = 8;
constexpr int kStDt = 0, kStAlpha = 1, kStSince = 2, kStDone = 3, kStStep = 4;
constexpr int kResRms = 0, kResFmax = 1, kResSteps = 2, kResStatus = 3;
constexpr int kStWords = 3 * kStSlot;
constexpr int kMeshRunning = -1, kMeshConverged = 0, kMeshNotConverged = 1, kMeshNonfinite = 2;
// Most workgroups of a point or element launch (grid-stride beyond).
constexpr int kMeshMaxBlocks = 2048;

struct MeshDev {
    int n;                       // points
    int nb, nt, nf;              // bonds, triples, facets
    double *x, *v, *f;           // 2n each
    // Each element writes its forces into its own slots, in element order
    // (contiguous writes); a point sums the slots that name it.
    const int2 *bond;            // (i, j)
    const double2 *bond_geo;     // (rest length, x image offset of j)
    const int4 *tri;             // (hinge, end a, end b, 0)
    const double2 *tri_off;      // x image offsets of the two arms
    const int4 *facet;           // (p1, p2, p3, 0)
    const double *inv;           // 6 per facet
    const double *area;          // per facet
    double2 *slot;               // element forces: bonds [0, nb), triple ends a then
                                 // b (nt each), facet corners 0, 1, 2 (nf each)
    // Slot entries of each point, (slot << 1) | negate, in element order, in
    // three sections (bonds, triples, facets), each stored column-major:
    // entry k of point i in section s at ent[sec[s] + k * n + i], so the
    // k-th entries of consecutive points are consecutive words.
    const int *ent;
    const int *cnt;              // n: entries per section, 8 bits each
    long sec[3];                 // section offsets in ent
    const int *twin;             // n: twin point or -1; null when twins are not coupled
    const unsigned char *fcode;  // n clamp codes
    double *part;                // 4 per gather workgroup
    double *tot;                 // 4: the partials' sums (v.F, v.v, F.F, max |F|^2)
    unsigned *ticket;            // gather workgroups done (0 between launches)
    double *st;                  // kStWords
    int gather_blocks;
    double ks, kb, smod, bmod, pref_d;
    double alpha0, f_alpha, dt_max, f_inc, f_dec, mass, fcut, sqrt_dof;
    int nmin, max_steps, criterion, pad;
};

int mesh_element_blocks(const MeshDev &m);
int mesh_point_blocks(int n);
// phase 0: one FIRE step reading state slot par; 1: the initial force and
// the first drift (reads slot par, writes par ^ 1).
int launch_mesh_step(const MeshDev &m, int par, int phase, void *stream);
int launch_mesh_affine(double *x, int n, double miny, double yrange, double xdisp, double ydisp,
                       void *stream);
int launch_mesh_energy(const MeshDev &m, double *part, int blocks, void *stream);

}  // namespace nr
