/*
 * CPU FP64 oracle for the network relaxation step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * code: only tests/, the smoke check in __graft_entry__.py and the CPU-baseline
 * leg of bench.py may use it, and only as the checker.
 *
 * It restates, in plain C, the algorithm of the reference's relax step
 * (src/large_strain_relax.cu) from the reference's behaviour, function by
 * function (each definition in oracle.c cites the reference lines it follows):
 *   read_network  :42-114   get_triples :116-184   clamp codes :896-902
 *   axial_forces  :386-414  bending_forces :417-469
 *   axial_energies :472-490 bending_energies :493-523 affine_displace :527-537
 *   run_fire_md   :596-762  do_relaxation_run :792-1010
 *   parse_batch_file :297-383, report_name :203-213, report_displaced :216-227
 *   get_file_lists (src/file_lists.cpp:11-137)
 * and, in linear.c, the small-strain solver of src/sparse_relax.cpp
 * (stiffness assembly :286-428, reduced system :560-605, solve and
 * energies :49-96, :676-881); in analysis.c, the reductions of
 * src/na_calc.cpp:308-326, src/strain_calc.cpp:96-120 and
 * src/axial_contribution.cpp:173-190; in mesh.c, the network + FEM-gel
 * composite of src/network_mesh_relax.cu.
 * plus the extensions the product also carries (Lees-Edwards box, max-force
 * criterion, per-fiber bending stiffness, nonlinear CG).
 *
 * Floating-point conventions shared with the product so that results can be
 * compared tightly: axpy-style updates are fma(); squared lengths are
 * fma(x, x, y*y); the rest length is sqrt(dx*dx + dy*dy) without contraction,
 * as the reference's host compiler produces it; the bending angle deviation is
 * atan2(-cross, -dot), which equals the reference's atan2(cross, dot) mapped to
 * [0, 2pi) minus pi.
 */
#ifndef NR_ORACLE_H
#define NR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n, e, t;             /* nodes, bonds, bending triples */
    double width;            /* x period (header "offset") */
    double height;           /* y period; 0 = reference mode (clamped rows) */
    double le_shift;         /* Lees-Edwards image shift in x (strain*height) */
    double miny, maxy;       /* extreme y of the input geometry */
    double *pos;             /* 2n, input geometry */
    int *bond;               /* 2e */
    double *rest;            /* e */
    signed char *bwrap;      /* 2e: wx, wy */
    int *tri;                /* 3t: hinge, end a, end b */
    signed char *twrap;      /* 4t: wx_a, wy_a, wx_b, wy_b */
    double *tkb;             /* t per-triple bending stiffness, or NULL */
    int *tbond;              /* 2t: the two bonds of each triple */
    unsigned char *fcode;    /* n: bit0 x free, bit1 y free */
} or_net;

/* error codes */
#define OR_OK 0
#define OR_E_OPEN 1
#define OR_E_HEADER 2
#define OR_E_POINTS 3
#define OR_E_INDEX 4
#define OR_E_EDGES 5
#define OR_E_ALLOC 6
#define OR_E_ARG 7
#define OR_E_FACETS 12  /* composite: facet or area section invalid */

int or_read_network(const char *path, or_net *net);
int or_net_from_arrays(int n, int e, double width, double height,
                       const double *pos, const int *bond,
                       const signed char *bwrap, or_net *net);
void or_free(or_net *net);
int or_get_triples(or_net *net);
void or_set_clamps(or_net *net);
int or_set_fiber_kappa(or_net *net, double klo, double khi, uint64_t seed);
int or_fiber_labels(const or_net *net, int *label);

void or_forces(const or_net *net, double ks, double kb, const double *pos,
               double *force);
void or_energies(const or_net *net, double ks, double kb, const double *pos,
                 double *es, double *eb);
void or_affine(const or_net *net, double *pos, double miny, double yrange,
               double xdisp, double ydisp);
/* Shear virial dE/dgamma (affine shear of the element geometry); the shear
 * stress is this over the box area.  Extension: no reference counterpart. */
double or_shear_virial(const or_net *net, double ks, double kb, const double *pos);

/* FIRE parameters: alpha0, f_alpha, dt_max, f_inc, f_dec */
#define OR_CRIT_RMS 0
#define OR_CRIT_MAX 1
/* report: [0] rms |F|, [1] E_stretch, [2] E_bend, [3] max node |F| */
int or_fire(const or_net *net, double ks, double kb, double *pos,
            const double fp[5], int nmin, double mass, double fcut,
            int max_steps, int criterion, double report[4], int *steps);
/* Nonlinear CG: Fletcher-Reeves (pr = 0) or Polak-Ribiere+ (pr != 0,
 * beta = max(0, F.(F - F_old) / |F_old|^2)). */
int or_cg(const or_net *net, double ks, double kb, double *pos, double fcut,
          int max_iters, int criterion, int pr, double report[4], int *iters);

/* Whole strain protocol with the reference's output files (strain_mode as
 * the reference's default, absolute displacements when 0). */
typedef struct {
    double ks, kb;
    double fp[5];
    int nmin;
    double mass, fcut;
    int nmax;
    double cmin, cmax, cinc;
    double smin, smax, sinc;
    int criterion;
    int method;              /* 0 FIRE, 1 CG */
    double kappa_lo, kappa_hi; /* per-fiber kappa when kappa_hi > 0 */
    uint64_t kappa_seed;
} or_batch;

int or_parse_batch(const char *path, char *base, int base_cap, char *tag,
                   int tag_cap, or_batch *b);
int or_run_protocol(const char *net_file, const char *log_name,
                    const char *edata_name, const char *disp_base,
                    const or_batch *b, int strain_mode);

int or_dec_digits(int value);
void or_report_name(const char *base, int num_digits, int count,
                    const char *ext, char *out, int cap);

/* Linear (small-strain) response, src/sparse_relax.cpp (oracle/linear.c).
 * K: the assembled stiffness, both triangles, CSR over the 2n DOFs. */
typedef struct {
    int n;        /* rows (2 x nodes) */
    long nnz;
    int *rowptr;  /* n + 1 */
    int *col;
    double *val;
} or_csr;
int or_linear_stiffness(const or_net *net, double ks, double kb, double mu, or_csr *K);
void or_csr_free(or_csr *K);
/* s, c: shear and compression strains of the affine field (affine != 0), or
 * the top row's displacements (affine == 0, reference mode only).
 * u: 2n displacements.  rep: [0] E_stretch+shear, [1] E_bend, [2] residual
 * 2-norm of the reduced system, [3] CG iterations. */
int or_linear_response(const or_net *net, double ks, double kb, double mu, double s, double c,
                       int affine, double tol, int max_iter, double *u, double rep[4]);
void or_linear_energies(const or_net *net, double ks, double kb, double mu, const double *u,
                        double s, double c, int affine, double *e_ss, double *e_bend);

/* Post-relax reductions (oracle/analysis.c): out = {non-affinity, axial
 * energy, mean bond strain, max |bond strain|}; strain: NULL or e values. */
void or_analyze(const or_net *net, const double *x0, const double *x, double shift,
                double shear, double axial, double k, double *strain, double out[4]);

/* Network + FEM-gel composite (oracle/mesh.c, src/network_mesh_relax.cu).
 * net holds the points (all of them), bonds, triples and clamp codes; its
 * width is the box offset, grown by sqrt(1 + dilation) at set-up. */
typedef struct {
    or_net net;
    int nf;              /* facets */
    int *facet;          /* 3 nf point indices */
    double *area;        /* nf, as read */
    double *inv;         /* 6 nf inverse matrices (undilated geometry) */
    int np;              /* twin pairs */
    int *pair;           /* 2 np: left-edge point, its twin one offset right */
    double offset;       /* header offset as read */
    double minx, maxx;   /* extreme x of the input geometry */
} or_mesh;

typedef struct {
    double ks, kb, bmod, smod, dilation;
    double fp[5];
    int nmin;
    double mass, fcut;
    int nmax;
    double cmin, cmax, cinc;
    double smin, smax, sinc;
} or_mesh_batch;

int or_read_mesh(const char *path, or_mesh *m);
void or_mesh_free(or_mesh *m);
void or_mesh_inv_mats(or_mesh *m, const double *pos);
int or_mesh_setup(or_mesh *m, double ks, double kb, double bmod, double dilation, double *pos,
                  double *pref_d);
double or_gel_energy(const or_mesh *m, double smod, double bmod, double pd, const double *pos);
void or_mesh_forces(const or_mesh *m, double ks, double kb, double smod, double bmod, double pd,
                    const double *pos, double *force);
void or_mesh_energies(const or_mesh *m, double ks, double kb, double smod, double bmod,
                      double pd, const double *pos, double e[3]);
/* report: [0] rms |F|, [1] E_stretch, [2] E_bend, [3] E_gel, [4] max |F| */
int or_mesh_fire(const or_mesh *m, double ks, double kb, double smod, double bmod, double pd,
                 double *pos, const double fp[5], int nmin, double mass, double fcut,
                 int max_steps, double report[5], int *steps);
int or_parse_mesh_batch(const char *path, char *base, int base_cap, char *tag, int tag_cap,
                        or_mesh_batch *b);
int or_run_mesh_protocol(const char *net_file, const char *log_name, const char *edata_name,
                         const char *disp_base, const or_mesh_batch *b, int strain_mode,
                         double *pref_d_out);

/* Synthetic triangular lattice (spacing 1, odd rows shifted by 1/2), each
 * candidate bond (node*3 + direction) kept when u < p. */
int or_lattice(int L, double p, uint64_t seed, int le_mode, or_net *net);
double or_uniform(uint64_t seed, uint64_t counter);

#ifdef __cplusplus
}
#endif
#endif
