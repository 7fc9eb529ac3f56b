/*
 * Post-relax reductions of a network state: a CPU FP64 restatement of the
 * reference's analysis programs.  TEST INFRASTRUCTURE ONLY: used by tests/
 * as the checker of the product's device reductions, never by the product.
 *
 *   calc_na_val  (src/na_calc.cpp:308-326): mean over nodes of
 *                |u - u_e|^2, u_e = ((y - miny) shear, (y - miny) axial),
 *                y, miny of the input geometry
 *   get_strains  (src/strain_calc.cpp:96-120): (l_f - l_i) / l_i per bond,
 *                d_f = d_i + u_j - u_i
 *   axial energy (src/axial_contribution.cpp:173-190): k/2 sum
 *                ((u_j - u_i) . rhat)^2, rhat = d_i / |d_i|
 * u = x - x0.  In a Lees-Edwards box a bond's second node seen across the
 * y boundary (wrap wy) is displaced by wy x shift in x in addition (an
 * extension: the reference's tools know only x-periodic networks).
 */
#include "oracle.h"

#include <math.h>

void or_analyze(const or_net *net, const double *x0, const double *x, double shift,
                double shear, double axial, double k, double *strain, double out[4])
{
    double na = 0.0, ax = 0.0, sum = 0.0, mx = 0.0;
    int i, e;
    for (i = 0; i < net->n; i++) {
        const double ux = x[2 * i] - x0[2 * i], uy = x[2 * i + 1] - x0[2 * i + 1];
        const double y = x0[2 * i + 1];
        const double ex = (y - net->miny) * shear, ey = (y - net->miny) * axial;
        na += (ux - ex) * (ux - ex) + (uy - ey) * (uy - ey);
    }
    for (e = 0; e < net->e; e++) {
        const int a = net->bond[2 * e], b = net->bond[2 * e + 1];
        const int wx = net->bwrap[2 * e], wy = net->bwrap[2 * e + 1];
        const double dxi = x0[2 * b] + wx * net->width - x0[2 * a];
        const double dyi = x0[2 * b + 1] + wy * net->height - x0[2 * a + 1];
        const double uxa = x[2 * a] - x0[2 * a], uya = x[2 * a + 1] - x0[2 * a + 1];
        const double uxb = x[2 * b] - x0[2 * b] + wy * shift, uyb = x[2 * b + 1] - x0[2 * b + 1];
        const double dxf = dxi + uxb - uxa, dyf = dyi + uyb - uya;
        const double li = sqrt(dxi * dxi + dyi * dyi), lf = sqrt(dxf * dxf + dyf * dyf);
        const double st = (lf - li) / li;
        if (strain) strain[e] = st;
        const double proj = (uxb - uxa) * (dxi / li) + (uyb - uya) * (dyi / li);
        ax += 0.5 * proj * proj;
        sum += st;
        mx = fmax(mx, fabs(st));
    }
    out[0] = net->n > 0 ? na / net->n : 0.0;
    out[1] = k * ax;
    out[2] = net->e > 0 ? sum / net->e : 0.0;
    out[3] = mx;
}
