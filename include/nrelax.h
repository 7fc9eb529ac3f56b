/*
 * nrelax.h -- C ABI of the MI355X network relaxation library (libnrelax.so).
 *
 * The library relaxes 2-D bond-stretch + fiber-bend networks under imposed
 * shear/compression by FIRE (and nonlinear CG) on one gfx950 GPU.  It is the
 * drop-in for the relax step of the reference program `lsrelax`
 * (src/large_strain_relax.cu).  Plain pointers and sizes only; every function
 * returns an int status (NR_OK == 0) unless stated otherwise, and
 * nr_last_error() gives a message for the calling thread's last failure.
 *
 * Handles are thread-compatible: use one handle from one thread at a time.
 */
#ifndef NRELAX_H
#define NRELAX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define NR_OK 0
#define NR_E_OPEN 1      /* file cannot be opened */
#define NR_E_HEADER 2    /* .dat header invalid (reference :60-65) */
#define NR_E_POINTS 3    /* too few points (reference :81-85) */
#define NR_E_INDEX 4     /* bond index out of range (reference :100-104) */
#define NR_E_EDGES 5     /* too few edges (reference :111) */
#define NR_E_ALLOC 6     /* host or device allocation failed */
#define NR_E_ARG 7       /* invalid argument */
#define NR_E_HIP 8       /* HIP runtime error (no GPU, launch failure, ...) */
#define NR_E_TOPO 9      /* topology the device layout cannot hold */
#define NR_E_BATCH 10    /* batch file invalid (reference :297-383) */
#define NR_E_TIMEOUT 11  /* a device wait exceeded its bound */
#define NR_E_MESH 12     /* composite: facet or area section invalid
                            (network_mesh_relax.cu:154-219) */

/* ---- per-relax outcome (nr_report.status) ----------------------------- */
#define NR_CONVERGED 0
#define NR_NOT_CONVERGED 1   /* step limit reached (logged, not an error) */
#define NR_NONFINITE 2       /* force norm became NaN/Inf */
#define NR_ABORTED 3         /* device wait bound exceeded */
#define NR_IN_PROGRESS 4     /* serving: step budget of a launch reached, the
                                relax continues in the next launch */

#define NR_CRIT_RMS 0        /* |F|_2 / sqrt(2N) < fcut (reference :716) */
#define NR_CRIT_MAX 1        /* max over nodes of |F_node| < fcut */

#define NR_METHOD_FIRE 0
#define NR_METHOD_CG 1       /* nonlinear CG, Fletcher-Reeves (the method the
                                reference names, large_strain_types.hpp:31-34) */
#define NR_METHOD_CG_PR 2    /* nonlinear CG, Polak-Ribiere with beta =
                                max(0, F.(F - F_old) / |F_old|^2): a restart
                                whenever that is not positive */

#define NR_PREC_FP64 0       /* FP64 throughout */
#define NR_PREC_MIXED 1      /* FP32 forces on FP64 anchor + FP32 displacement,
                                FP64 reductions, FP64 polish */

typedef struct nr_network nr_network;  /* host-side network + topology */
typedef struct nr_batch nr_batch;      /* device-resident set of networks */

typedef struct {
    int n_nodes, n_bonds, n_triples;
    double width;      /* x period (reference .dat header "offset") */
    double height;     /* y period; 0 = reference mode (top/bottom clamped) */
    double miny, maxy; /* extreme y of the input geometry */
} nr_network_info;

/* FIRE / CG parameters: the reference's batch lines 3 and 4
 * (large_strain_relax.cu:334-354), plus criterion and method. */
typedef struct {
    double alpha0, f_alpha, dt_max, f_inc, f_dec; /* fire_params[5] */
    int nmin;
    double mass;
    double fcut;
    int max_steps;
    int criterion;     /* NR_CRIT_RMS or NR_CRIT_MAX */
    int method;        /* NR_METHOD_FIRE, NR_METHOD_CG or NR_METHOD_CG_PR */
} nr_relax_params;

/* Result of one relax: report_data[3] and step_count of run_fire_md
 * (large_strain_relax.cu:596, :721, :743, :751) plus max force and status. */
typedef struct {
    double rms_force;  /* |F|_2 / sqrt(2N) */
    double max_force;  /* max node |F| */
    double e_stretch;  /* sum 1/2 ks (d - L)^2 */
    double e_bend;     /* sum 1/2 kb (theta - pi)^2 */
    int steps;         /* steps taken, the final one included */
    int status;        /* NR_CONVERGED, NR_NOT_CONVERGED, ... */
    int fp64_steps;    /* mixed precision: steps of the FP64 polish */
    int pad;
    double t_start;    /* device clock (s) when this network's relax began */
    double t_done;     /* device clock (s) when it finished */
    double shear_stress; /* sigma_xy = (dE/dgamma of an affine shear) / box
                            area, box = width x height (Lees-Edwards) or
                            width x (maxy - miny); an extension, the
                            reference reports no stress */
} nr_report;

typedef struct {
    int precision;     /* NR_PREC_FP64 / NR_PREC_MIXED */
    int block;         /* threads per workgroup, 0 = tuned */
    int tile_nodes;    /* nodes per tile, 0 = tuned */
    int team;          /* most workgroups per network, 0 = tuned */
    int device;        /* HIP device ordinal */
    int tiles_per_member; /* a network of T tiles gets a team of about
                             T / tiles_per_member workgroups.  > 0: team
                             sizes depend on each network alone, so its
                             results do not depend on the rest of the batch;
                             0 = tuned for the batch (1, 2, 4 or 8) */
    double mixed_switch;  /* NR_PREC_MIXED: the FP32 stage hands over to FP64
                             when max |F| < mixed_switch x its initial value,
                             0 = default (3e-2) */
} nr_batch_opts;

/* ---- errors ------------------------------------------------------------ */
const char *nr_last_error(void);
const char *nr_version(void);

/* ---- host-side networks (no GPU needed) -------------------------------- */

/* read_network (large_strain_relax.cu:42-114): header "N E width [height]",
 * N point lines "x y", E edge lines "i j wx [wy]"; rest lengths from the
 * input geometry.  A 4th header number selects the Lees-Edwards box. */
int nr_network_read(const char *path, nr_network **out);
int nr_network_from_arrays(int n_nodes, int n_bonds, double width,
                           double height, const double *pos, const int *bonds,
                           const signed char *wraps, nr_network **out);
/* Seeded triangular lattice L x L (spacing 1), bond kept with probability p. */
int nr_network_lattice(int L, double p, uint64_t seed, int lees_edwards,
                       nr_network **out);
/* get_triples (large_strain_relax.cu:116-184). */
int nr_network_build_triples(nr_network *net);
/* Per-fiber bending stiffness, log-uniform in [lo, hi] from seed. */
int nr_network_set_fiber_kappa(nr_network *net, double lo, double hi,
                               uint64_t seed);
int nr_network_get_info(const nr_network *net, nr_network_info *info);
/* Copies out any non-NULL array: pos[2N], bonds[2E], bond_wraps[2E],
 * rest[E], triples[3T], triple_wraps[4T], triple_kappa[T] (zeros when
 * uniform), fcodes[N] (clamp codes, reference :896-902). */
int nr_network_get_arrays(const nr_network *net, double *pos, int *bonds,
                          signed char *bond_wraps, double *rest, int *triples,
                          signed char *triple_wraps, double *triple_kappa,
                          unsigned char *fcodes);
int nr_network_write_dat(const nr_network *net, const char *path);
void nr_network_free(nr_network *net);

/* ---- device batches ----------------------------------------------------- */

/* Uploads `count` networks (triples must already be built when kb > 0) with
 * stiffnesses ks[i], kb[i].  Positions start at each network's input
 * geometry.  opts may be NULL (FP64, tuned shapes, device 0). */
int nr_batch_create(int count, nr_network *const *nets, const double *ks,
                    const double *kb, const nr_batch_opts *opts,
                    nr_batch **out);
/* Affine step of affine_displace (large_strain_relax.cu:527-537) for one
 * network (index) or all (index < 0): x += xdisp (y - miny)/yrange,
 * y += ydisp (y - miny)/yrange, applied at the start of the next relax, fused
 * into its first force pass.  yrange = maxy - miny (reference mode) or the box
 * height (Lees-Edwards, where xdisp also advances the image shift). */
int nr_batch_strain(nr_batch *b, int index, double xdisp, double ydisp);
/* Relaxes every network of the batch (run_fire_md, :596-762) in one
 * persistent launch; params is one struct per network, or one shared struct
 * when shared != 0.  reports[count] receives the outcomes. */
int nr_batch_relax(nr_batch *b, const nr_relax_params *params, int shared,
                   nr_report *reports);
/* Strain ramp of do_relaxation_run (large_strain_relax.cu:917-1000, the
 * compression then the shear loop) for every network in ONE launch: the
 * networks take the n_inc affine steps incs[2i] (xdisp), incs[2i+1] (ydisp)
 * in turn, each relaxed (run_fire_md) from the previous relaxed state, with
 * no host round trip between increments.  reports receives count * n_inc
 * outcomes, network-major; a network whose relax turns non-finite stops its
 * ramp there (its later reports are left untouched). */
int nr_batch_ramp(nr_batch *b, int n_inc, const double *incs,
                  const nr_relax_params *params, int shared, nr_report *reports);
/* One strain step of do_relaxation_run's loops (large_strain_relax.cu:
 * 925-931, 967-973: affine_displace then run_fire_md) captured as a HIP
 * graph: network k takes the affine step (xdisp[k], ydisp[k]) on the device,
 * then every network relaxes with params.  nr_batch_graph_replay runs the
 * whole step as one graph launch (the host's network records are uploaded
 * first, so steps added with nr_batch_strain and positions set in between
 * are honoured).  Capturing again replaces the graph.  The graph keeps the
 * team plan (class allocation and queue order) of the batch's last relax:
 * capture after a relax that was itself planned from known step counts (the
 * second relax of the batch onwards) to replay with the plan of a warm relax.
 * A network relaxed by a team of several workgroups needs the whole team
 * resident at once: eager relaxes launch such teams cooperatively, but a
 * graph node is an ordinary launch whose grid only fits the device's resident
 * workgroups.  Replays therefore need the GPU to themselves; when another
 * process's kernels hold workgroup slots, a team's bounded waits end the
 * replay with NR_E_TIMEOUT (after 60 s) instead of hanging. */
int nr_batch_graph_capture(nr_batch *b, const double *xdisp, const double *ydisp,
                           const nr_relax_params *params, int shared);
int nr_batch_graph_replay(nr_batch *b, nr_report *reports);
/* Restores every network to its input geometry (zero strain, no pending
 * step) without host traffic. */
int nr_batch_reset(nr_batch *b);
int nr_batch_get_positions(nr_batch *b, int index, double *pos);
int nr_batch_set_positions(nr_batch *b, int index, const double *pos);
/* One force / energy evaluation at the current positions (pending strain
 * applied first), for checking: force[2N], energies[2] = {stretch, bend}. */
int nr_batch_eval(nr_batch *b, int index, double *force, double *energies);
/* Post-relax reductions of network `index` at its current state, computed on
 * the device from the displacements u = x - x0 since the batch was created
 * (or reset):
 *   nonaffinity      mean over nodes of |u - u_aff|^2 with u_aff =
 *                    ((y0 - miny) shear_strain, (y0 - miny) axial_strain):
 *                    calc_na_val, na_calc.cpp:308-326
 *   axial_energy     k_axial/2 sum over bonds of (u_ij . rhat_ij)^2, rhat from
 *                    the input geometry: axial_contribution.cpp:173-190
 *   mean_bond_strain, max_bond_strain (max |.|) of the bond strains
 *                    (l - l0) / l0: get_strains, strain_calc.cpp:96-120
 * bond_strains (NULL or one double per bond) receives the strains in the
 * network's edge order.  In a Lees-Edwards box the image of a bond's second
 * node across the y boundary moves with the box shift.  The bond list is read
 * from the network given to nr_batch_create, which must still exist (the
 * first call per network uploads it; later calls reuse the upload). */
typedef struct {
    double nonaffinity;
    double axial_energy;
    double mean_bond_strain;
    double max_bond_strain;
} nr_analysis;
int nr_batch_analyze(nr_batch *b, int index, double shear_strain, double axial_strain,
                     double k_axial, nr_analysis *out, double *bond_strains);
/* No counterpart in the reference: nr_batch_analyze for callers that cannot
 * allocate memory (code running while a HIP graph is captured, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results, bit
 * for bit; its partial sums, bond strains and (when no nr_batch_analyze call
 * uploaded it before) the network's bond list live in a workspace the library
 * holds on each GPU it runs on until the process ends (24 MiB + 8 KiB: at
 * most NR_NOALLOC_MAX_BONDS bonds, NR_E_ARG beyond that), so the call makes
 * no device or host allocation.  Every call on one GPU shares that workspace:
 * calls that overlap on one GPU (other threads, other batches) race. */
#define NR_NOALLOC_MAX_BONDS (1 << 20)
int nr_batch_analyze_noalloc(nr_batch *b, int index, double shear_strain, double axial_strain,
                             double k_axial, nr_analysis *out, double *bond_strains);
/* Device time of the last relax launch in milliseconds. */
double nr_batch_last_ms(const nr_batch *b);

/* ---- Small-strain linear response (sparse_relax.cpp) ------------------- */
/* Replaces do_relaxation_run of sparse_relax.cpp (:676-881) for one load step:
 * the stiffness assembled from the reference's element blocks (:286-428),
 * the reduced system over the free DOFs K_rr w = -K_rN bc (get_lhs_rhs
 * :560-605) solved on the current HIP device by conjugate gradients from
 * w = 0 (the reference factorises it by sparse QR; CG gives the
 * minimum-norm solution: floppy modes and bond-less nodes do not move), to
 * |K_rr w + K_rN bc| <= tol x the imposed field's force scale or max_iter
 * iterations.  bc: with affine != 0 the affine field of shear and
 * compression strains (affine_displacement :527-537), else the top row
 * displaced by (shear, -compression) (add_displacement :513-523; reference
 * mode only).  In a Lees-Edwards box every node is free and images across
 * the y boundary carry the field's offset.  u (NULL or 2N doubles) receives
 * bc + w; the report the linearised energies (get_stretch_shear_energy
 * :49-70 with gel modulus mu, get_bending_energy :73-96) and the residual
 * 2-norm (the reference's "Residual RMS force" line, :841-850).  Bending
 * (kb > 0 or per-fiber kappa) needs nr_network_build_triples first. */
typedef struct {
    double e_stretch;   /* stretch + gel shear energy */
    double e_bend;
    double residual;
    int iterations;
    int status;         /* NR_CONVERGED or NR_NOT_CONVERGED */
    double device_ms;   /* device time of the iterations */
} nr_linear_report;
int nr_linear_response(const nr_network *net, double ks, double kb, double mu, double shear,
                       double compression, int affine, double tol, int max_iter, double *u,
                       nr_linear_report *rep);
/* Several load steps, of one network or several (the reference solves them
 * one after another in do_relaxation_run :676-881 and run_relaxation_batch
 * :1040-1090), solved at once: every CG launch serves all of them, each with
 * its own scalars, stopping rule and workgroups, and the systems of one
 * network and stiffness share its matrix.  Each job's results equal, bit for
 * bit, those of nr_linear_response with the same arguments (alone or in any
 * batch); device_ms is the whole batch's device time. */
typedef struct {
    const nr_network *net;
    double ks, kb, mu, shear, compression;
    int affine;
    double tol;
    int max_iter;
    double *u;          /* NULL or 2N doubles: bc + w */
} nr_linear_job;
int nr_linear_response_batch(int count, const nr_linear_job *jobs, nr_linear_report *reps);
/* The small-strain batch of the reference program `sp_relax`
 * (run_relaxation_batch, sparse_relax.cpp:1040-1090, and do_relaxation_run
 * :676-881): the 4-line batch file (base [tag]; ks kb mu; compression min max
 * inc; shear min max inc), every <base>*.dat, one device solve per load step
 * (nr_linear_response to 1e-12 of the force scale), and the reference's
 * outputs: <m>_<tag>_log.txt ("Residual RMS force: %2.10le" per step, then the
 * elapsed time), <m>_<tag>_edata.txt ("%12.8lf\t%12.8lf\t%12.10le" compression,
 * shear, energy; itemize: stretch and bend energies apart) and
 * <m>_<tag>_<NNN>.disp (the displacement field).  strain_mode (-s) scales the
 * compression by the y extent and the shear by the x extent, affine = 0 (-z)
 * displaces the top row only.  The reference's null-space removal (-n) has
 * nothing to do here: the CG solution carries no null-space component. */
int nr_linear_batch(const char *batch_file, int strain_mode, int affine, int itemize);
/* Bytes one FIRE step moves per the algorithmic count for this batch. */
int nr_batch_counts(const nr_batch *b, int index, int64_t *nodes,
                    int64_t *bonds, int64_t *triples);
void nr_batch_destroy(nr_batch *b);

/* Shape tuner (no counterpart in the reference, whose grid and block come
 * from the -x/-y flags, large_strain_relax.cu:1059-1063): for these networks
 * picks the workgroup/tile shape and the tiles-per-member rule (1, 2, 4 or 8,
 * or opts->tiles_per_member when > 0) by timing each candidate with a probe
 * relax: network k takes a step budget proportional to its expected step
 * count (a fit over network size and bond density; the budgets average
 * probe_steps), with no convergence test.  opts is read (precision, device,
 * team) and written with the fastest choice; *predicted_ms (may be NULL) is
 * its probe time scaled to the expected step counts.  Results depend on the
 * shape only through reduction order, so use one choice for every batch
 * whose results are compared. */
int nr_autotune(int count, nr_network *const *nets, const double *ks,
                const double *kb, int probe_steps, nr_batch_opts *opts,
                double *predicted_ms);

/* ---- the reference program's relax protocol ---------------------------- */

/* relaxation_batch + do_relaxation_run (large_strain_relax.cu:792-1044):
 * parses the 6-line batch file, finds <base>*.dat, and for each network runs
 * the compression then shear ramps, writing <m>_<tag>_log.txt,
 * <m>_<tag>_edata.txt and <m>_<tag>_<NNN>.pos as the reference does.
 * strain_mode != 0 scales displacements by the y range (the default). */
int nr_relaxation_batch(const char *batch_file, int strain_mode,
                        const nr_batch_opts *opts);

/* ---- network + FEM-gel composite (src/network_mesh_relax.cu) ----------- */

/* A composite is a bond/bend network on the points of a triangular gel mesh
 * (.nmc file: header "points edges facets offset", points, edges
 * "i j wrap", facets "i j k", one area per facet).  The gel is periodic in x
 * through twin points: each left-edge point with a point one offset to its
 * right (within 1e-10); twins receive the sum of their forces. */
typedef struct nr_mesh nr_mesh;
typedef struct nr_composite nr_composite;

typedef struct {
    int n_points, n_bonds, n_facets, n_twins;
    double offset;      /* header offset (the x period) */
    double minx, maxx, miny, maxy;
} nr_mesh_info;

/* read_network_mesh (network_mesh_relax.cu:72-230) with twin pairs
 * (:219-226).  Errors as nr_network_read, NR_E_MESH for the facet and area
 * sections. */
int nr_mesh_read(const char *path, nr_mesh **out);
/* Synthetic composite (no reference counterpart): the reference-mode L x L
 * lattice of nr_network_lattice on a mesh of the strip 0 <= x <= L, with
 * mesh-only points on the strip's edges and each boundary half-cell as a
 * right triangle. */
int nr_mesh_lattice(int L, double p, uint64_t seed, nr_mesh **out);
int nr_mesh_write_nmc(const nr_mesh *m, const char *path);
int nr_mesh_get_info(const nr_mesh *m, nr_mesh_info *info);
/* Copies (any may be NULL): pos 2 x points, facets 3 x facets, areas,
 * twins 2 x twin pairs. */
int nr_mesh_get_arrays(const nr_mesh *m, double *pos, int *facets, double *areas, int *twins);
void nr_mesh_free(nr_mesh *m);

/* Line 2 of the composite batch file (parse_batch_file :433-440): fiber
 * stiffnesses, the gel's bulk and shear moduli (gel_energy :674-699:
 * A [smod (|F|^2/J - 2) + bmod (J - 1 - d)^2] / 2 per facet) and the
 * pre-swelling dilation. */
typedef struct {
    double ks, kb, bmod, smod, dilation;
} nr_gel_params;

typedef struct {
    double width;       /* offset x sqrt(1 + dilation), the bonds' x period */
    double range;       /* (maxy - miny) x sqrt(1 + dilation) */
    double pref_d;      /* the gel's preferred dilation d (bmod + net_bmod) / bmod */
    double net_bmod;    /* 2 E_stretch / (area dilation^2) of the dilated network */
    int n_points, n_bonds, n_triples, n_facets; /* elements on the device */
} nr_composite_info;

/* The set-up of do_relaxation_run (:1138-1275) on a copy of m, resident on
 * device `device`: triples (kb > 0) found with the dilated offset on the
 * undilated points (as the reference does), clamp codes and the facets'
 * inverse matrices (get_inv_mats :503-531) from the undilated points, the
 * points dilated by sqrt(1 + dilation) about (min x, min y) (dilate
 * :796-806) and the preferred dilation (:1256-1275; 0 when dilation is 0,
 * where the reference divides 0 by 0).  Terms whose modulus is 0 are left
 * out, as the reference skips their kernels.  FIRE steps run as replays of
 * a captured HIP graph; NRELAX_GRAPHS=0 in the environment at creation
 * launches the same kernels directly instead. */
int nr_composite_create(const nr_mesh *m, const nr_gel_params *g, int device,
                        nr_composite **out);
int nr_composite_get_info(const nr_composite *c, nr_composite_info *info);
/* affine_displace (:783-793): x += xdisp (y - min y) / yrange, y likewise. */
int nr_composite_affine(nr_composite *c, double yrange, double xdisp, double ydisp);
/* run_fire_md (:865-1066) from the resident positions: FIRE in FP64 (method
 * must be NR_METHOD_FIRE; criterion as nr_relax_params).  rep as for
 * networks (e_stretch, e_bend, rms/max force, steps, status); *e_gel (may be
 * NULL) the gel energy. */
int nr_composite_relax(nr_composite *c, const nr_relax_params *p, nr_report *rep,
                       double *e_gel);
int nr_composite_get_positions(nr_composite *c, double *pos);
int nr_composite_set_positions(nr_composite *c, const double *pos);
/* The clamped force with twins summed (force: NULL or 2 x points) and
 * {E_stretch, E_bend, E_gel} (energies: NULL or 3) at the resident
 * positions, by the relax's own kernels. */
int nr_composite_eval(nr_composite *c, double *force, double *energies);
/* Device time of the last relax (ms). */
double nr_composite_last_ms(const nr_composite *c);
void nr_composite_destroy(nr_composite *c);
/* relaxation_batch + do_relaxation_run of the composite program
 * (:1095-1421): the 6-line batch file (line 2 "ks kb bmod [smod
 * [dilation]]"), every <base>*.nmc, "Preferred dilation: " on stdout, the
 * compression ramp from its minimum (the height shrinking with each step),
 * then the shear ramp, and the reference's outputs: <m>_<tag>_log.txt,
 * <m>_<tag>_edata.txt (compression "%12.0lf\t%12.10lf" + four energies,
 * shear "%12.8lf\t%12.8lf" + stretch, bend, gel and total "%12.8le") and
 * <m>_<tag>_<NNN>.pos after each shear step. */
int nr_mesh_relaxation_batch(const char *batch_file, int strain_mode, int device);

/* ---- the XML job format (src/large_strain_types.hpp) ------------------- */

/* The reference declares an XML relaxation job (element names, children and
 * attributes :36-50, RelaxScheme / Run / RelaxBatch :99-168, the description
 * :226-250) but never parses it (parse_relaxation_batch :266-273 is empty).
 * <rbatch> holds <name input tag>, <mech_params stretch bend>, <rscheme
 * type="FIRE|ConjGrad" pars="..."> and one or more <run> with <axial min max
 * inc>, <shear min max inc> and <report edata log disp>.  FIRE pars are the
 * batch file's lines 3 and 4 (alpha0 f_alpha dt_max f_inc f_dec N_min mass
 * fcut max_steps); ConjGrad pars "fcut max_iters" or the same nine.  Each
 * <base>*.dat runs its runs in order on its resident state, as the batch
 * protocol (nr_relaxation_batch) does one ramp pair; one run writes the batch
 * protocol's files, several write run k's as <m>_<tag>_run<k>_... ; a false
 * report flag skips that file.  NR_E_BATCH for a file that breaks the
 * format (nr_last_error says where). */
int nr_xml_batch(const char *job_file, int strain_mode, const nr_batch_opts *opts);
/* Parses and checks a job file and lists what nr_xml_batch would do (the
 * job's settings, then one line per network and run with its output names)
 * into plan (NUL-terminated, at most cap bytes).  No GPU work. */
int nr_xml_job_plan(const char *job_file, char *plan, int cap);

/* ---- serving ------------------------------------------------------------ */

/* A request loop with continuous batching on one GPU; no counterpart in the
 * reference, which relaxes one network at a time.  Jobs are laid out on host
 * worker threads, admitted as device memory and slots free up (arrival order;
 * smaller jobs fill in behind one that does not fit yet, until it has waited
 * a second), and relaxed by repeated launches of the persistent kernel that
 * give every resident job at most slice_steps FIRE steps.  Each job's state
 * stays on the device across launches and across its strain increments.  A
 * job's arithmetic does not depend on the other resident jobs, so its result
 * is the same alone and in any mix.  One thread drives a server. */
typedef struct nr_server nr_server;

typedef struct {
    int64_t hbm_bytes;    /* device arena for resident jobs, 0 = 8 GiB */
    int max_jobs;         /* resident job slots, 0 = 4096 */
    int slice_steps;      /* FIRE steps per job per launch, 0 = 2000 */
    int planner_threads;  /* host threads building device layouts, 0 = 4 */
} nr_server_opts;

typedef struct {
    int job;              /* id from nr_server_submit */
    int status;           /* NR_CONVERGED when every increment converged, else
                             the first other outcome (NR_ABORTED with `error`
                             set when the job could not be laid out or placed) */
    int error;            /* NR_OK or an NR_E_ code */
    int n_inc;            /* increments with a report */
    int steps;            /* FIRE steps over all increments */
    int n_nodes;
    double t_submit, t_admit, t_done;     /* nr_server_clock() seconds */
    double e_stretch, e_bend, max_force;  /* after the last increment */
    /* the post-relax reductions of nr_batch_analyze at the job's final state,
     * computed on the device when the job retires: non-affinity against the
     * affine field of the applied increments (sum of xdisp and of ydisp over
     * the y range), axial energy with k = ks, mean and max |bond strain| */
    double nonaffinity, axial_energy, mean_bond_strain, max_bond_strain;
} nr_job_result;

typedef struct {
    int64_t slices;       /* launches so far */
    double busy_ms;       /* device time of those launches */
    int64_t node_cap, nodes_free;
    int resident_workgroups;
    int active, waiting, done;
} nr_server_info;

/* shape: precision, block, tile_nodes, team, device, tiles_per_member,
 * mixed_switch as for nr_batch_create (NULL = defaults; the server's default
 * is one tile per team member, the widest teams, for latency). */
int nr_server_create(const nr_batch_opts *shape, const nr_server_opts *opts,
                     nr_server **out);
/* Seconds since the server was created (the clock of nr_job_result). */
double nr_server_clock(const nr_server *s);
/* Queues a job: the network is copied (triples built when kb > 0); it takes
 * the n_inc affine steps incs[2i] (xdisp), incs[2i+1] (ydisp) in turn, each
 * relaxed from the previous relaxed state (the loop of do_relaxation_run,
 * large_strain_relax.cu:917-1000).  keep_positions != 0 keeps the final
 * positions for nr_server_result.  *job receives the id. */
int nr_server_submit(nr_server *s, const nr_network *net, double ks, double kb,
                     const nr_relax_params *params, int n_inc, const double *incs,
                     int keep_positions, int *job);
/* One scheduling round: waits for the launch in flight, retires the jobs it
 * finished, starts the next launch over the resident jobs, then admits
 * waiting jobs (their uploads overlap that launch; they join the next one).
 * Returns once the launch is started. */
int nr_server_step(nr_server *s, int *n_active, int *n_waiting);
/* Steps until every submitted job is finished. */
int nr_server_drain(nr_server *s);
/* NR_OK and the id of the oldest finished job not yet taken, or 1 if none. */
int nr_server_next_done(nr_server *s, int *job);
/* Copies a finished job's result, its n_inc reports (reports may be NULL)
 * and, when kept, its final positions pos[2N] (may be NULL), then forgets
 * the job. */
int nr_server_result(nr_server *s, int job, nr_job_result *res, nr_report *reports,
                     double *pos);
int nr_server_stats(const nr_server *s, nr_server_info *st);
void nr_server_destroy(nr_server *s);

#ifdef __cplusplus
}
#endif
#endif
