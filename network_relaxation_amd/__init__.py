"""MI355X network relaxation: Python view of libnrelax.so's C ABI.

The product is the C ABI in include/nrelax.h (libnrelax.so, HIP kernels for
gfx950) and the `nrelax` CLI, a drop-in for the reference's `lsrelax -b`.
This module binds the C ABI with ctypes for tests and benchmarks.  Nothing
here computes on the CPU: every force, energy and relax runs in the library,
and importing raises if the library has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnrelax.so")
CLI_PATH = os.path.join(_HERE, "nrelax")
LINEAR_CLI_PATH = os.path.join(_HERE, "nrelax_linear")
MESH_CLI_PATH = os.path.join(_HERE, "nrelax_mesh")

OK = 0
(E_OPEN, E_HEADER, E_POINTS, E_INDEX, E_EDGES, E_ALLOC, E_ARG, E_HIP, E_TOPO, E_BATCH, E_TIMEOUT,
 E_MESH) = range(1, 13)
CONVERGED, NOT_CONVERGED, NONFINITE, ABORTED, IN_PROGRESS = range(5)
CRIT_RMS, CRIT_MAX = 0, 1
METHOD_FIRE, METHOD_CG, METHOD_CG_PR = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1


class NetworkInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_bonds", C.c_int), ("n_triples", C.c_int),
                ("width", C.c_double), ("height", C.c_double),
                ("miny", C.c_double), ("maxy", C.c_double)]


class RelaxParams(C.Structure):
    _fields_ = [("alpha0", C.c_double), ("f_alpha", C.c_double), ("dt_max", C.c_double),
                ("f_inc", C.c_double), ("f_dec", C.c_double), ("nmin", C.c_int),
                ("mass", C.c_double), ("fcut", C.c_double), ("max_steps", C.c_int),
                ("criterion", C.c_int), ("method", C.c_int)]


class Report(C.Structure):
    _fields_ = [("rms_force", C.c_double), ("max_force", C.c_double),
                ("e_stretch", C.c_double), ("e_bend", C.c_double),
                ("steps", C.c_int), ("status", C.c_int), ("fp64_steps", C.c_int),
                ("pad", C.c_int), ("t_start", C.c_double), ("t_done", C.c_double),
                ("shear_stress", C.c_double)]

    def as_dict(self):
        return dict(rms=self.rms_force, fmax=self.max_force, es=self.e_stretch,
                    eb=self.e_bend, steps=self.steps, status=self.status,
                    fp64_steps=self.fp64_steps, t_start=self.t_start, t_done=self.t_done,
                    sxy=self.shear_stress)


class BatchOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("block", C.c_int), ("tile_nodes", C.c_int),
                ("team", C.c_int), ("device", C.c_int), ("tiles_per_member", C.c_int),
                ("mixed_switch", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_bytes", C.c_int64), ("max_jobs", C.c_int), ("slice_steps", C.c_int),
                ("planner_threads", C.c_int)]


class JobResult(C.Structure):
    _fields_ = [("job", C.c_int), ("status", C.c_int), ("error", C.c_int), ("n_inc", C.c_int),
                ("steps", C.c_int), ("n_nodes", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("e_stretch", C.c_double),
                ("e_bend", C.c_double), ("max_force", C.c_double),
                ("nonaffinity", C.c_double), ("axial_energy", C.c_double),
                ("mean_bond_strain", C.c_double), ("max_bond_strain", C.c_double)]


class ServerInfo(C.Structure):
    _fields_ = [("slices", C.c_int64), ("busy_ms", C.c_double), ("node_cap", C.c_int64),
                ("nodes_free", C.c_int64), ("resident_workgroups", C.c_int),
                ("active", C.c_int), ("waiting", C.c_int), ("done", C.c_int)]


class Analysis(C.Structure):
    _fields_ = [("nonaffinity", C.c_double), ("axial_energy", C.c_double),
                ("mean_bond_strain", C.c_double), ("max_bond_strain", C.c_double)]


class LinearReport(C.Structure):
    _fields_ = [("e_stretch", C.c_double), ("e_bend", C.c_double), ("residual", C.c_double),
                ("iterations", C.c_int), ("status", C.c_int), ("device_ms", C.c_double)]


class LinearJob(C.Structure):
    _fields_ = [("net", C.c_void_p), ("ks", C.c_double), ("kb", C.c_double), ("mu", C.c_double),
                ("shear", C.c_double), ("compression", C.c_double), ("affine", C.c_int),
                ("tol", C.c_double), ("max_iter", C.c_int), ("u", C.POINTER(C.c_double))]


class MeshInfo(C.Structure):
    _fields_ = [("n_points", C.c_int), ("n_bonds", C.c_int), ("n_facets", C.c_int),
                ("n_twins", C.c_int), ("offset", C.c_double), ("minx", C.c_double),
                ("maxx", C.c_double), ("miny", C.c_double), ("maxy", C.c_double)]


class GelParams(C.Structure):
    _fields_ = [("ks", C.c_double), ("kb", C.c_double), ("bmod", C.c_double),
                ("smod", C.c_double), ("dilation", C.c_double)]


class CompositeInfo(C.Structure):
    _fields_ = [("width", C.c_double), ("range", C.c_double), ("pref_d", C.c_double),
                ("net_bmod", C.c_double), ("n_points", C.c_int), ("n_bonds", C.c_int),
                ("n_triples", C.c_int), ("n_facets", C.c_int)]


class NRError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nrelax error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Loads libnrelax.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build with `make -C network_relaxation_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        P, vp, dp, ip = C.POINTER, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.nr_last_error.restype = C.c_char_p
        L.nr_version.restype = C.c_char_p
        L.nr_network_read.argtypes = [C.c_char_p, P(vp)]
        L.nr_network_from_arrays.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, dp, ip,
                                             P(C.c_byte), P(vp)]
        L.nr_network_lattice.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_int, P(vp)]
        L.nr_network_build_triples.argtypes = [vp]
        L.nr_network_set_fiber_kappa.argtypes = [vp, C.c_double, C.c_double, C.c_uint64]
        L.nr_network_get_info.argtypes = [vp, P(NetworkInfo)]
        L.nr_network_get_arrays.argtypes = [vp, dp, ip, P(C.c_byte), dp, ip, P(C.c_byte), dp,
                                            P(C.c_ubyte)]
        L.nr_network_write_dat.argtypes = [vp, C.c_char_p]
        L.nr_network_free.argtypes = [vp]
        L.nr_network_free.restype = None
        L.nr_batch_create.argtypes = [C.c_int, P(vp), dp, dp, P(BatchOpts), P(vp)]
        L.nr_batch_strain.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.nr_batch_relax.argtypes = [vp, P(RelaxParams), C.c_int, P(Report)]
        L.nr_batch_reset.argtypes = [vp]
        L.nr_batch_get_positions.argtypes = [vp, C.c_int, dp]
        L.nr_batch_set_positions.argtypes = [vp, C.c_int, dp]
        L.nr_batch_eval.argtypes = [vp, C.c_int, dp, dp]
        L.nr_batch_analyze.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                       P(Analysis), dp]
        L.nr_batch_analyze_noalloc.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                               P(Analysis), dp]
        L.nr_linear_response.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_int, C.c_double, C.c_int, dp,
                                         P(LinearReport)]
        L.nr_linear_response_batch.argtypes = [C.c_int, P(LinearJob), P(LinearReport)]
        L.nr_batch_last_ms.argtypes = [vp]
        L.nr_batch_last_ms.restype = C.c_double
        L.nr_batch_counts.argtypes = [vp, C.c_int, P(C.c_int64), P(C.c_int64), P(C.c_int64)]
        L.nr_batch_destroy.argtypes = [vp]
        L.nr_batch_destroy.restype = None
        L.nr_relaxation_batch.argtypes = [C.c_char_p, C.c_int, P(BatchOpts)]
        L.nr_batch_ramp.argtypes = [vp, C.c_int, dp, P(RelaxParams), C.c_int, P(Report)]
        L.nr_batch_graph_capture.argtypes = [vp, dp, dp, P(RelaxParams), C.c_int]
        L.nr_batch_graph_replay.argtypes = [vp, P(Report)]
        L.nr_autotune.argtypes = [C.c_int, P(vp), dp, dp, C.c_int, P(BatchOpts), dp]
        L.nr_server_create.argtypes = [P(BatchOpts), P(ServerOpts), P(vp)]
        L.nr_server_clock.argtypes = [vp]
        L.nr_server_clock.restype = C.c_double
        L.nr_server_submit.argtypes = [vp, vp, C.c_double, C.c_double, P(RelaxParams), C.c_int,
                                       dp, C.c_int, ip]
        L.nr_server_step.argtypes = [vp, ip, ip]
        L.nr_server_drain.argtypes = [vp]
        L.nr_server_next_done.argtypes = [vp, ip]
        L.nr_server_result.argtypes = [vp, C.c_int, P(JobResult), P(Report), dp]
        L.nr_server_stats.argtypes = [vp, P(ServerInfo)]
        L.nr_server_destroy.argtypes = [vp]
        L.nr_mesh_read.argtypes = [C.c_char_p, P(vp)]
        L.nr_mesh_lattice.argtypes = [C.c_int, C.c_double, C.c_uint64, P(vp)]
        L.nr_mesh_write_nmc.argtypes = [vp, C.c_char_p]
        L.nr_mesh_get_info.argtypes = [vp, P(MeshInfo)]
        L.nr_mesh_get_arrays.argtypes = [vp, dp, ip, dp, ip]
        L.nr_mesh_free.argtypes = [vp]
        L.nr_mesh_free.restype = None
        L.nr_composite_create.argtypes = [vp, P(GelParams), C.c_int, P(vp)]
        L.nr_composite_get_info.argtypes = [vp, P(CompositeInfo)]
        L.nr_composite_affine.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.nr_composite_relax.argtypes = [vp, P(RelaxParams), P(Report), dp]
        L.nr_composite_get_positions.argtypes = [vp, dp]
        L.nr_composite_set_positions.argtypes = [vp, dp]
        L.nr_composite_eval.argtypes = [vp, dp, dp]
        L.nr_composite_last_ms.argtypes = [vp]
        L.nr_composite_last_ms.restype = C.c_double
        L.nr_composite_destroy.argtypes = [vp]
        L.nr_composite_destroy.restype = None
        L.nr_mesh_relaxation_batch.argtypes = [C.c_char_p, C.c_int, C.c_int]
        L.nr_xml_batch.argtypes = [C.c_char_p, C.c_int, P(BatchOpts)]
        L.nr_xml_job_plan.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise NRError(rc, lib().nr_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Network:
    """Host-side network (nr_network): .dat reader, lattice, triples."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        _check(lib().nr_network_read(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    @classmethod
    def lattice(cls, L, p, seed, lees_edwards=True):
        h = C.c_void_p()
        _check(lib().nr_network_lattice(L, p, seed, 1 if lees_edwards else 0, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_arrays(cls, pos, bonds, wraps, width, height=0.0):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
        bonds = np.ascontiguousarray(bonds, np.int32).reshape(-1)
        wraps = np.ascontiguousarray(wraps, np.int8).reshape(-1)
        h = C.c_void_p()
        _check(lib().nr_network_from_arrays(len(pos) // 2, len(bonds) // 2, width, height,
                                            _dp(pos), bonds.ctypes.data_as(C.POINTER(C.c_int)),
                                            wraps.ctypes.data_as(C.POINTER(C.c_byte)),
                                            C.byref(h)))
        return cls(h.value)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.nr_network_free(self.h)
            self.h = None

    def triples(self):
        _check(lib().nr_network_build_triples(self.h))
        return self

    def fiber_kappa(self, lo, hi, seed):
        _check(lib().nr_network_set_fiber_kappa(self.h, lo, hi, seed))
        return self

    def info(self):
        i = NetworkInfo()
        _check(lib().nr_network_get_info(self.h, C.byref(i)))
        return i

    def arrays(self):
        i = self.info()
        n, e, t = i.n_nodes, i.n_bonds, i.n_triples
        pos = np.zeros(2 * n)
        bonds = np.zeros(2 * max(e, 1), np.int32)
        bw = np.zeros(2 * max(e, 1), np.int8)
        rest = np.zeros(max(e, 1))
        tri = np.zeros(3 * max(t, 1), np.int32)
        tw = np.zeros(4 * max(t, 1), np.int8)
        tk = np.zeros(max(t, 1))
        fc = np.zeros(n, np.uint8)
        _check(lib().nr_network_get_arrays(
            self.h, _dp(pos), bonds.ctypes.data_as(C.POINTER(C.c_int)),
            bw.ctypes.data_as(C.POINTER(C.c_byte)), _dp(rest),
            tri.ctypes.data_as(C.POINTER(C.c_int)), tw.ctypes.data_as(C.POINTER(C.c_byte)),
            _dp(tk), fc.ctypes.data_as(C.POINTER(C.c_ubyte))))
        return dict(pos=pos, bonds=bonds[:2 * e].reshape(-1, 2), bond_wraps=bw[:2 * e].reshape(-1, 2),
                    rest=rest[:e], triples=tri[:3 * t].reshape(-1, 3),
                    triple_wraps=tw[:4 * t].reshape(-1, 4), triple_kappa=tk[:t], fcodes=fc)

    def write_dat(self, path):
        _check(lib().nr_network_write_dat(self.h, os.fsencode(path)))


def relax_params(alpha0=0.1, f_alpha=0.99, dt_max=0.1, f_inc=1.1, f_dec=0.5, nmin=5, mass=1.0,
                 fcut=1e-10, max_steps=1000000, criterion=CRIT_MAX, method=METHOD_FIRE):
    return RelaxParams(alpha0, f_alpha, dt_max, f_inc, f_dec, nmin, mass, fcut, max_steps,
                       criterion, method)


class Batch:
    """Device-resident networks (nr_batch) relaxed by one persistent launch."""

    def __init__(self, networks, ks, kb, precision=PREC_FP64, tile_nodes=0, team=0, device=0,
                 block=0, tiles_per_member=0, mixed_switch=0.0):
        self.networks = list(networks)  # keep the host networks alive
        n = len(self.networks)
        ks = np.broadcast_to(np.asarray(ks, np.float64), (n,)).copy()
        kb = np.broadcast_to(np.asarray(kb, np.float64), (n,)).copy()
        handles = (C.c_void_p * n)(*[x.h.value for x in self.networks])
        opts = BatchOpts(precision, block, tile_nodes, team, device, tiles_per_member,
                         mixed_switch)
        h = C.c_void_p()
        _check(lib().nr_batch_create(n, handles, _dp(ks), _dp(kb), C.byref(opts), C.byref(h)))
        self.h = h
        self.count = n
        self.n_nodes = [x.info().n_nodes for x in self.networks]

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.nr_batch_destroy(self.h)
            self.h = None

    def strain(self, index, xdisp, ydisp=0.0):
        _check(lib().nr_batch_strain(self.h, index, xdisp, ydisp))

    def shear(self, index, gamma):
        """Shear strain increment gamma (strain mode: displacement gamma*yrange)."""
        idx = range(self.count) if index < 0 else [index]
        for k in idx:
            i = self.networks[k].info()
            yr = i.height if i.height > 0 else i.maxy - i.miny
            _check(lib().nr_batch_strain(self.h, k, gamma * yr, 0.0))

    def strain_all(self, xdisp, ydisp=None):
        """Per-network affine steps (nr_batch_strain for each network)."""
        for k in range(self.count):
            _check(lib().nr_batch_strain(self.h, k, float(xdisp[k]),
                                         0.0 if ydisp is None else float(ydisp[k])))

    def relax(self, params):
        reps = (Report * self.count)()
        if isinstance(params, RelaxParams):
            _check(lib().nr_batch_relax(self.h, C.byref(params), 1, reps))
        else:
            arr = (RelaxParams * self.count)(*params)
            _check(lib().nr_batch_relax(self.h, arr, 0, reps))
        return [r.as_dict() for r in reps]

    def ramp(self, incs, params):
        """nr_batch_ramp: incs is (n_inc, 2) of (xdisp, ydisp) for every network;
        returns reports[count][n_inc]."""
        incs = np.ascontiguousarray(incs, np.float64).reshape(-1, 2)
        n = len(incs)
        reps = (Report * (self.count * n))()
        if isinstance(params, RelaxParams):
            _check(lib().nr_batch_ramp(self.h, n, _dp(incs), C.byref(params), 1, reps))
        else:
            arr = (RelaxParams * self.count)(*params)
            _check(lib().nr_batch_ramp(self.h, n, _dp(incs), arr, 0, reps))
        return [[reps[k * n + i].as_dict() for i in range(n)] for k in range(self.count)]

    def graph_capture(self, xdisp, ydisp, params):
        """nr_batch_graph_capture: one strain step (per-network affine steps)
        plus the relax, recorded as a HIP graph."""
        xd = np.ascontiguousarray(np.broadcast_to(np.asarray(xdisp, np.float64), (self.count,)))
        yd = np.ascontiguousarray(np.broadcast_to(np.asarray(ydisp, np.float64), (self.count,)))
        if isinstance(params, RelaxParams):
            _check(lib().nr_batch_graph_capture(self.h, _dp(xd), _dp(yd), C.byref(params), 1))
        else:
            arr = (RelaxParams * self.count)(*params)
            _check(lib().nr_batch_graph_capture(self.h, _dp(xd), _dp(yd), arr, 0))

    def graph_replay(self):
        reps = (Report * self.count)()
        _check(lib().nr_batch_graph_replay(self.h, reps))
        return [r.as_dict() for r in reps]

    def reset(self):
        _check(lib().nr_batch_reset(self.h))

    def positions(self, index):
        out = np.zeros(2 * self.n_nodes[index])
        _check(lib().nr_batch_get_positions(self.h, index, _dp(out)))
        return out

    def set_positions(self, index, pos):
        pos = np.ascontiguousarray(pos, np.float64)
        _check(lib().nr_batch_set_positions(self.h, index, _dp(pos)))

    def eval(self, index):
        f = np.zeros(2 * self.n_nodes[index])
        e = np.zeros(2)
        _check(lib().nr_batch_eval(self.h, index, _dp(f), _dp(e)))
        return f, e

    def analyze(self, index, shear, axial=0.0, k_axial=1.0, strains=False, noalloc=False):
        """Post-relax reductions of network `index` on the device
        (nr_batch_analyze): non-affinity against the affine field of `shear`
        and `axial`, axial energy, mean / max |bond strain|, and with
        strains=True the bond strains in edge order.  noalloc=True calls
        nr_batch_analyze_noalloc (the library's fixed workspace)."""
        a = Analysis()
        st = np.zeros(max(self.networks[index].info().n_bonds, 1)) if strains else None
        fn = lib().nr_batch_analyze_noalloc if noalloc else lib().nr_batch_analyze
        _check(fn(self.h, index, shear, axial, k_axial, C.byref(a),
                  _dp(st) if strains else None))
        out = dict(nonaffinity=a.nonaffinity, axial_energy=a.axial_energy,
                   mean_strain=a.mean_bond_strain, max_strain=a.max_bond_strain)
        if strains:
            out["strains"] = st[:self.networks[index].info().n_bonds]
        return out

    def last_ms(self):
        return lib().nr_batch_last_ms(self.h)

    def counts(self, index=-1):
        n, e, t = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().nr_batch_counts(self.h, index, C.byref(n), C.byref(e), C.byref(t)))
        return n.value, e.value, t.value


class Server:
    """nr_server: continuous batching of relax jobs on one GPU."""

    def __init__(self, precision=PREC_FP64, tile_nodes=0, team=0, device=0, block=0,
                 tiles_per_member=0, mixed_switch=0.0, hbm_bytes=0, max_jobs=0, slice_steps=0,
                 planner_threads=0):
        shape = BatchOpts(precision, block, tile_nodes, team, device, tiles_per_member,
                          mixed_switch)
        so = ServerOpts(hbm_bytes, max_jobs, slice_steps, planner_threads)
        h = C.c_void_p()
        _check(lib().nr_server_create(C.byref(shape), C.byref(so), C.byref(h)))
        self.h = h
        self.meta = {}

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.nr_server_destroy(self.h)
            self.h = None

    def clock(self):
        return lib().nr_server_clock(self.h)

    def submit(self, net, ks, kb, params, incs, keep_positions=False):
        """Queues `net` for the strain increments incs ((n, 2) of xdisp, ydisp)."""
        incs = np.ascontiguousarray(incs, np.float64).reshape(-1, 2)
        j = C.c_int()
        _check(lib().nr_server_submit(self.h, net.h, ks, kb, C.byref(params), len(incs),
                                      _dp(incs), 1 if keep_positions else 0, C.byref(j)))
        self.meta[j.value] = (net.info().n_nodes, len(incs), keep_positions)
        return j.value

    def step(self):
        a, w = C.c_int(), C.c_int()
        _check(lib().nr_server_step(self.h, C.byref(a), C.byref(w)))
        return a.value, w.value

    def drain(self):
        _check(lib().nr_server_drain(self.h))

    def results(self):
        """Takes every finished job's result (list of dicts)."""
        out = []
        j = C.c_int()
        while True:
            rc = lib().nr_server_next_done(self.h, C.byref(j))
            if rc == 1:
                return out
            _check(rc)
            n, ninc, keep = self.meta.pop(j.value)
            res = JobResult()
            reps = (Report * ninc)()
            pos = np.zeros(2 * n) if keep else None
            _check(lib().nr_server_result(self.h, j.value, C.byref(res), reps,
                                          _dp(pos) if keep else None))
            d = {f: getattr(res, f) for f, _ in JobResult._fields_}
            d["reports"] = [reps[i].as_dict() for i in range(res.n_inc)]
            d["pos"] = pos
            out.append(d)

    def stats(self):
        i = ServerInfo()
        _check(lib().nr_server_stats(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in ServerInfo._fields_}


def autotune(networks, ks, kb, precision=PREC_FP64, probe_steps=200, device=0, team=0):
    """nr_autotune: returns (shape kwargs for Batch/Server, predicted ms)."""
    n = len(networks)
    ks = np.broadcast_to(np.asarray(ks, np.float64), (n,)).copy()
    kb = np.broadcast_to(np.asarray(kb, np.float64), (n,)).copy()
    handles = (C.c_void_p * n)(*[x.h.value for x in networks])
    opts = BatchOpts(precision, 0, 0, team, device, 0, 0.0)
    pred = C.c_double()
    _check(lib().nr_autotune(n, handles, _dp(ks), _dp(kb), probe_steps, C.byref(opts),
                             C.byref(pred)))
    return dict(block=opts.block, tile_nodes=opts.tile_nodes,
                tiles_per_member=opts.tiles_per_member), pred.value


def relaxation_batch(batch_file, strain_mode=True, precision=PREC_FP64, tile_nodes=0,
                     tiles_per_member=0):
    """nr_relaxation_batch: the reference program's whole protocol, every
    matched network in one device batch."""
    opts = BatchOpts(precision, 0, tile_nodes, 0, 0, tiles_per_member, 0.0)
    return lib().nr_relaxation_batch(os.fsencode(batch_file), 1 if strain_mode else 0,
                                     C.byref(opts))


def linear_response(net, ks, kb, shear, compression=0.0, mu=0.0, affine=True, tol=1e-12,
                    max_iter=1000000, displacements=False):
    """Small-strain linear response of one network on the current device
    (nr_linear_response): linearised energies, residual, CG iterations, and
    with displacements=True the displacement field u (2N)."""
    rep = LinearReport()
    u = np.zeros(2 * net.info().n_nodes) if displacements else None
    _check(lib().nr_linear_response(net.h, ks, kb, mu, shear, compression, int(affine), tol,
                                    max_iter, _dp(u) if displacements else None, C.byref(rep)))
    out = dict(es=rep.e_stretch, eb=rep.e_bend, residual=rep.residual,
               iters=rep.iterations, status=rep.status, device_ms=rep.device_ms)
    if displacements:
        out["u"] = u
    return out


def linear_response_batch(jobs, displacements=False):
    """Several small-strain solves at once on the current device
    (nr_linear_response_batch).  jobs: dicts with net, ks, kb, shear and
    optionally compression, mu, affine, tol, max_iter (linear_response's
    defaults).  Returns one result per job, as linear_response does; each
    equals linear_response's for the same arguments bit for bit."""
    n = len(jobs)
    arr = (LinearJob * max(n, 1))()
    us = []
    for k, j in enumerate(jobs):
        u = np.zeros(2 * j["net"].info().n_nodes) if displacements else None
        us.append(u)
        arr[k] = LinearJob(j["net"].h, j["ks"], j["kb"], j.get("mu", 0.0), j["shear"],
                           j.get("compression", 0.0), int(j.get("affine", True)),
                           j.get("tol", 1e-12), j.get("max_iter", 1000000),
                           _dp(u) if displacements else None)
    reps = (LinearReport * max(n, 1))()
    _check(lib().nr_linear_response_batch(n, arr, reps))
    out = []
    for k in range(n):
        r = reps[k]
        d = dict(es=r.e_stretch, eb=r.e_bend, residual=r.residual, iters=r.iterations,
                 status=r.status, device_ms=r.device_ms)
        if displacements:
            d["u"] = us[k]
        out.append(d)
    return out


class Mesh:
    """Host-side network + gel composite (nr_mesh): .nmc reader, lattice."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)

    @classmethod
    def read(cls, path):
        h = C.c_void_p()
        _check(lib().nr_mesh_read(os.fsencode(path), C.byref(h)))
        return cls(h.value)

    @classmethod
    def lattice(cls, L, p, seed):
        h = C.c_void_p()
        _check(lib().nr_mesh_lattice(L, p, seed, C.byref(h)))
        return cls(h.value)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.nr_mesh_free(self.h)
            self.h = None

    def info(self):
        i = MeshInfo()
        _check(lib().nr_mesh_get_info(self.h, C.byref(i)))
        return i

    def arrays(self):
        i = self.info()
        pos = np.zeros(2 * i.n_points)
        fac = np.zeros(3 * max(i.n_facets, 1), np.int32)
        area = np.zeros(max(i.n_facets, 1))
        tw = np.zeros(2 * max(i.n_twins, 1), np.int32)
        ip = C.POINTER(C.c_int)
        _check(lib().nr_mesh_get_arrays(self.h, _dp(pos), fac.ctypes.data_as(ip), _dp(area),
                                        tw.ctypes.data_as(ip)))
        return dict(pos=pos, facets=fac[:3 * i.n_facets].reshape(-1, 3),
                    areas=area[:i.n_facets], twins=tw[:2 * i.n_twins].reshape(-1, 2))

    def write_nmc(self, path):
        _check(lib().nr_mesh_write_nmc(self.h, os.fsencode(path)))


class Composite:
    """A composite resident on one GPU (nr_composite): set-up, affine steps,
    FIRE relax, positions, force and energies."""

    def __init__(self, mesh, ks, kb, bmod, smod=0.0, dilation=0.0, device=0):
        self.mesh = mesh
        g = GelParams(ks, kb, bmod, smod, dilation)
        h = C.c_void_p()
        _check(lib().nr_composite_create(mesh.h, C.byref(g), device, C.byref(h)))
        self.h = h
        self.n = mesh.info().n_points

    def __del__(self):
        if getattr(self, "h", None) and self.h.value and _lib is not None:
            _lib.nr_composite_destroy(self.h)
            self.h = None

    def info(self):
        i = CompositeInfo()
        _check(lib().nr_composite_get_info(self.h, C.byref(i)))
        return i

    def affine(self, yrange, xdisp, ydisp):
        _check(lib().nr_composite_affine(self.h, yrange, xdisp, ydisp))

    def relax(self, params):
        rep = Report()
        eg = C.c_double()
        _check(lib().nr_composite_relax(self.h, C.byref(params), C.byref(rep), C.byref(eg)))
        return dict(status=rep.status, steps=rep.steps, rms=rep.rms_force, fmax=rep.max_force,
                    es=rep.e_stretch, eb=rep.e_bend, eg=eg.value,
                    ms=lib().nr_composite_last_ms(self.h))

    def positions(self):
        pos = np.zeros(2 * self.n)
        _check(lib().nr_composite_get_positions(self.h, _dp(pos)))
        return pos

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, np.float64)
        _check(lib().nr_composite_set_positions(self.h, _dp(pos)))

    def eval(self):
        f = np.zeros(2 * self.n)
        e = np.zeros(3)
        _check(lib().nr_composite_eval(self.h, _dp(f), _dp(e)))
        return f, e


def mesh_relaxation_batch(batch_file, strain_mode=True, device=0):
    """nr_mesh_relaxation_batch: the composite program's whole protocol."""
    return lib().nr_mesh_relaxation_batch(os.fsencode(batch_file), 1 if strain_mode else 0,
                                          device)


def xml_job_plan(job_file):
    """nr_xml_job_plan: parse and check an XML job file; the plan as text
    (no GPU work).  Raises NRError for a file that breaks the format."""
    buf = C.create_string_buffer(1 << 16)
    _check(lib().nr_xml_job_plan(os.fsencode(job_file), buf, len(buf)))
    return buf.value.decode()


def xml_batch(job_file, strain_mode=True, precision=PREC_FP64):
    """nr_xml_batch: run an XML job (several runs per network)."""
    opts = BatchOpts(precision, 0, 0, 0, 0, 0, 0.0)
    return lib().nr_xml_batch(os.fsencode(job_file), 1 if strain_mode else 0, C.byref(opts))
