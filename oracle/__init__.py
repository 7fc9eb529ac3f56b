"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by
bench.py's CPU-baseline leg, as the checker.  The product never imports it.
The C sources (oracle.c, linear.c, analysis.c, mesh.c, oracle.h) cite the
reference lines they restate.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

CRIT_RMS = 0
CRIT_MAX = 1


class OrNet(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("e", C.c_int), ("t", C.c_int),
        ("width", C.c_double), ("height", C.c_double), ("le_shift", C.c_double),
        ("miny", C.c_double), ("maxy", C.c_double),
        ("pos", C.POINTER(C.c_double)),
        ("bond", C.POINTER(C.c_int)),
        ("rest", C.POINTER(C.c_double)),
        ("bwrap", C.POINTER(C.c_byte)),
        ("tri", C.POINTER(C.c_int)),
        ("twrap", C.POINTER(C.c_byte)),
        ("tkb", C.POINTER(C.c_double)),
        ("tbond", C.POINTER(C.c_int)),
        ("fcode", C.POINTER(C.c_ubyte)),
    ]


class OrBatch(C.Structure):
    _fields_ = [
        ("ks", C.c_double), ("kb", C.c_double), ("fp", C.c_double * 5),
        ("nmin", C.c_int), ("mass", C.c_double), ("fcut", C.c_double),
        ("nmax", C.c_int),
        ("cmin", C.c_double), ("cmax", C.c_double), ("cinc", C.c_double),
        ("smin", C.c_double), ("smax", C.c_double), ("sinc", C.c_double),
        ("criterion", C.c_int), ("method", C.c_int),
        ("kappa_lo", C.c_double), ("kappa_hi", C.c_double),
        ("kappa_seed", C.c_uint64),
    ]


class OrCsr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_long), ("rowptr", C.POINTER(C.c_int)),
                ("col", C.POINTER(C.c_int)), ("val", C.POINTER(C.c_double))]


class OrMesh(C.Structure):
    _fields_ = [
        ("net", OrNet), ("nf", C.c_int), ("facet", C.POINTER(C.c_int)),
        ("area", C.POINTER(C.c_double)), ("inv", C.POINTER(C.c_double)),
        ("np", C.c_int), ("pair", C.POINTER(C.c_int)),
        ("offset", C.c_double), ("minx", C.c_double), ("maxx", C.c_double),
    ]


class OrMeshBatch(C.Structure):
    _fields_ = [
        ("ks", C.c_double), ("kb", C.c_double), ("bmod", C.c_double), ("smod", C.c_double),
        ("dilation", C.c_double), ("fp", C.c_double * 5),
        ("nmin", C.c_int), ("mass", C.c_double), ("fcut", C.c_double), ("nmax", C.c_int),
        ("cmin", C.c_double), ("cmax", C.c_double), ("cinc", C.c_double),
        ("smin", C.c_double), ("smax", C.c_double), ("sinc", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.or_read_network.argtypes = [C.c_char_p, P(OrNet)]
        L.or_net_from_arrays.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double,
                                         dp, P(C.c_int), P(C.c_byte), P(OrNet)]
        L.or_free.argtypes = [P(OrNet)]
        L.or_free.restype = None
        L.or_get_triples.argtypes = [P(OrNet)]
        L.or_set_clamps.argtypes = [P(OrNet)]
        L.or_set_clamps.restype = None
        L.or_set_fiber_kappa.argtypes = [P(OrNet), C.c_double, C.c_double, C.c_uint64]
        L.or_fiber_labels.argtypes = [P(OrNet), P(C.c_int)]
        L.or_forces.argtypes = [P(OrNet), C.c_double, C.c_double, dp, dp]
        L.or_forces.restype = None
        L.or_energies.argtypes = [P(OrNet), C.c_double, C.c_double, dp, dp, dp]
        L.or_energies.restype = None
        L.or_affine.argtypes = [P(OrNet), dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_affine.restype = None
        L.or_fire.argtypes = [P(OrNet), C.c_double, C.c_double, dp, dp, C.c_int,
                              C.c_double, C.c_double, C.c_int, C.c_int, dp, P(C.c_int)]
        L.or_cg.argtypes = [P(OrNet), C.c_double, C.c_double, dp, C.c_double, C.c_int,
                            C.c_int, C.c_int, dp, P(C.c_int)]
        L.or_parse_batch.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p,
                                     C.c_int, P(OrBatch)]
        L.or_run_protocol.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                      P(OrBatch), C.c_int]
        L.or_dec_digits.argtypes = [C.c_int]
        L.or_report_name.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_char_p,
                                     C.c_char_p, C.c_int]
        L.or_report_name.restype = None
        L.or_lattice.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_int, P(OrNet)]
        L.or_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.or_uniform.restype = C.c_double
        L.or_linear_stiffness.argtypes = [P(OrNet), C.c_double, C.c_double, C.c_double,
                                          P(OrCsr)]
        L.or_csr_free.argtypes = [P(OrCsr)]
        L.or_csr_free.restype = None
        L.or_linear_response.argtypes = [P(OrNet), C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_int, C.c_double,
                                         C.c_int, dp, dp]
        L.or_linear_energies.argtypes = [P(OrNet), C.c_double, C.c_double, C.c_double, dp,
                                         C.c_double, C.c_double, C.c_int, dp, dp]
        L.or_linear_energies.restype = None
        L.or_analyze.argtypes = [P(OrNet), dp, dp, C.c_double, C.c_double, C.c_double,
                                 C.c_double, dp, dp]
        L.or_analyze.restype = None
        M = P(OrMesh)
        d = C.c_double
        L.or_read_mesh.argtypes = [C.c_char_p, M]
        L.or_mesh_free.argtypes = [M]
        L.or_mesh_free.restype = None
        L.or_mesh_inv_mats.argtypes = [M, dp]
        L.or_mesh_inv_mats.restype = None
        L.or_mesh_setup.argtypes = [M, d, d, d, d, dp, dp]
        L.or_gel_energy.argtypes = [M, d, d, d, dp]
        L.or_gel_energy.restype = d
        L.or_mesh_forces.argtypes = [M, d, d, d, d, d, dp, dp]
        L.or_mesh_forces.restype = None
        L.or_mesh_energies.argtypes = [M, d, d, d, d, d, dp, dp]
        L.or_mesh_energies.restype = None
        L.or_mesh_fire.argtypes = [M, d, d, d, d, d, dp, dp, C.c_int, d, d, C.c_int, dp,
                                   P(C.c_int)]
        L.or_parse_mesh_batch.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p,
                                          C.c_int, P(OrMeshBatch)]
        L.or_run_mesh_protocol.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                           P(OrMeshBatch), C.c_int, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Net:
    """Owns one or_net."""

    def __init__(self):
        self.s = OrNet()
        self._owned = False

    @classmethod
    def read(cls, path):
        n = cls()
        rc = lib().or_read_network(os.fsencode(path), C.byref(n.s))
        if rc:
            raise ValueError(f"or_read_network({path}) failed with code {rc}")
        n._owned = True
        return n

    @classmethod
    def lattice(cls, L, p, seed, le_mode=True):
        n = cls()
        rc = lib().or_lattice(L, p, seed, 1 if le_mode else 0, C.byref(n.s))
        if rc:
            raise ValueError(f"or_lattice failed with code {rc}")
        n._owned = True
        return n

    @classmethod
    def from_arrays(cls, pos, bonds, wraps, width, height=0.0):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        bonds = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1)
        wraps = np.ascontiguousarray(wraps, dtype=np.int8).reshape(-1)
        n = cls()
        rc = lib().or_net_from_arrays(len(pos) // 2, len(bonds) // 2, width, height,
                                      _dp(pos), bonds.ctypes.data_as(C.POINTER(C.c_int)),
                                      wraps.ctypes.data_as(C.POINTER(C.c_byte)),
                                      C.byref(n.s))
        if rc:
            raise ValueError(f"or_net_from_arrays failed with code {rc}")
        n._owned = True
        return n

    def __del__(self):
        if getattr(self, "_owned", False) and _lib is not None:
            _lib.or_free(C.byref(self.s))
            self._owned = False

    def triples(self):
        rc = lib().or_get_triples(C.byref(self.s))
        if rc:
            raise ValueError(f"or_get_triples failed with code {rc}")
        return self

    def fiber_kappa(self, klo, khi, seed):
        rc = lib().or_set_fiber_kappa(C.byref(self.s), klo, khi, seed)
        if rc:
            raise ValueError(f"or_set_fiber_kappa failed with code {rc}")
        return self

    # views (copies) of the arrays
    @property
    def n(self):
        return self.s.n

    @property
    def e(self):
        return self.s.e

    @property
    def t(self):
        return self.s.t

    def pos(self):
        return np.ctypeslib.as_array(self.s.pos, shape=(2 * self.s.n,)).copy()

    def bonds(self):
        return np.ctypeslib.as_array(self.s.bond, shape=(2 * self.s.e,)).reshape(-1, 2).copy()

    def bond_wraps(self):
        return np.ctypeslib.as_array(self.s.bwrap, shape=(2 * self.s.e,)).reshape(-1, 2).copy()

    def rest(self):
        return np.ctypeslib.as_array(self.s.rest, shape=(self.s.e,)).copy()

    def tris(self):
        if self.s.t == 0:
            return np.zeros((0, 3), np.int32)
        return np.ctypeslib.as_array(self.s.tri, shape=(3 * self.s.t,)).reshape(-1, 3).copy()

    def tri_wraps(self):
        if self.s.t == 0:
            return np.zeros((0, 4), np.int8)
        return np.ctypeslib.as_array(self.s.twrap, shape=(4 * self.s.t,)).reshape(-1, 4).copy()

    def tri_kappa(self):
        if self.s.t == 0 or not self.s.tkb:
            return None
        return np.ctypeslib.as_array(self.s.tkb, shape=(self.s.t,)).copy()

    def fiber_labels(self):
        lab = np.zeros(self.s.t, np.int32)
        rc = lib().or_fiber_labels(C.byref(self.s), lab.ctypes.data_as(C.POINTER(C.c_int)))
        if rc:
            raise ValueError("or_fiber_labels failed")
        return lab

    def fcode(self):
        return np.ctypeslib.as_array(self.s.fcode, shape=(self.s.n,)).copy()

    def forces(self, ks, kb, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        f = np.zeros(2 * self.s.n)
        lib().or_forces(C.byref(self.s), ks, kb, _dp(pos), _dp(f))
        return f

    def energies(self, ks, kb, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        es, eb = C.c_double(), C.c_double()
        lib().or_energies(C.byref(self.s), ks, kb, _dp(pos), C.byref(es), C.byref(eb))
        return es.value, eb.value

    def shear_virial(self, ks, kb, pos):
        """dE/dgamma under an affine shear of the element geometry."""
        f = lib().or_shear_virial
        f.restype = C.c_double
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        return f(C.byref(self.s), C.c_double(ks), C.c_double(kb), _dp(pos))

    def affine(self, pos, xdisp, ydisp, miny=None, yrange=None):
        """In-place affine step of affine_displace; in Lees-Edwards mode a shear
        step also advances the image shift."""
        if miny is None:
            miny = self.s.miny
        if yrange is None:
            yrange = self.s.height if self.s.height > 0 else self.s.maxy - self.s.miny
        lib().or_affine(C.byref(self.s), _dp(pos), miny, yrange, xdisp, ydisp)
        if self.s.height > 0:
            self.s.le_shift += xdisp

    def shear(self, pos, gamma):
        """Apply a shear strain increment gamma in strain mode."""
        yrange = self.s.height if self.s.height > 0 else self.s.maxy - self.s.miny
        self.affine(pos, gamma * yrange, 0.0)

    def fire(self, ks, kb, pos, fp=(0.1, 0.99, 0.1, 1.1, 0.5), nmin=5, mass=1.0,
             fcut=1e-10, max_steps=1000000, criterion=CRIT_MAX):
        fpa = (C.c_double * 5)(*fp)
        rep = (C.c_double * 4)()
        steps = C.c_int()
        rc = lib().or_fire(C.byref(self.s), ks, kb, _dp(pos), fpa, nmin, mass, fcut,
                           max_steps, criterion, rep, C.byref(steps))
        return dict(rc=rc, rms=rep[0], es=rep[1], eb=rep[2], fmax=rep[3], steps=steps.value)

    def linear_stiffness(self, ks, kb, mu=0.0):
        """Assembled stiffness (sparse_relax.cpp:286-428) as a dense 2n x 2n
        array (small networks only)."""
        K = OrCsr()
        if lib().or_linear_stiffness(C.byref(self.s), ks, kb, mu, C.byref(K)):
            raise MemoryError("or_linear_stiffness failed")
        try:
            n = K.n
            rp = np.ctypeslib.as_array(K.rowptr, shape=(n + 1,))
            col = np.ctypeslib.as_array(K.col, shape=(max(K.nnz, 1),))[:K.nnz]
            val = np.ctypeslib.as_array(K.val, shape=(max(K.nnz, 1),))[:K.nnz]
            D = np.zeros((n, n))
            for r in range(n):
                D[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
        finally:
            lib().or_csr_free(C.byref(K))
        return D

    def linear_response(self, ks, kb, s, c=0.0, mu=0.0, affine=True, tol=1e-12,
                        max_iter=1000000):
        """Small-strain response (sparse_relax.cpp:676-881): displacements u,
        linearised energies, the reduced system's residual, CG iterations."""
        u = np.zeros(2 * self.s.n)
        rep = np.zeros(4)
        rc = lib().or_linear_response(C.byref(self.s), ks, kb, mu, s, c, int(affine), tol,
                                      max_iter, _dp(u), _dp(rep))
        return dict(rc=rc, u=u, es=rep[0], eb=rep[1], residual=rep[2], iters=int(rep[3]))

    def linear_energies(self, ks, kb, u, s, c=0.0, mu=0.0, affine=True):
        u = np.ascontiguousarray(u, dtype=np.float64)
        es, eb = C.c_double(), C.c_double()
        lib().or_linear_energies(C.byref(self.s), ks, kb, mu, _dp(u), s, c, int(affine),
                                 C.byref(es), C.byref(eb))
        return es.value, eb.value

    def analyze(self, x0, x, shift, shear, axial=0.0, k=1.0):
        """Post-relax reductions (na_calc, strain_calc, axial_contribution):
        non-affinity, axial energy, mean and max |bond strain|, bond strains."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        st = np.zeros(max(self.s.e, 1))
        out = np.zeros(4)
        lib().or_analyze(C.byref(self.s), _dp(x0), _dp(x), shift, shear, axial, k, _dp(st),
                         _dp(out))
        return dict(nonaffinity=out[0], axial_energy=out[1], mean_strain=out[2],
                    max_strain=out[3], strains=st[:self.s.e])

    def cg(self, ks, kb, pos, fcut=1e-10, max_iters=1000000, criterion=CRIT_MAX, pr=False):
        """Nonlinear CG: Fletcher-Reeves, or Polak-Ribiere+ with pr=True."""
        rep = (C.c_double * 4)()
        it = C.c_int()
        rc = lib().or_cg(C.byref(self.s), ks, kb, _dp(pos), fcut, max_iters, criterion,
                         1 if pr else 0, rep, C.byref(it))
        return dict(rc=rc, rms=rep[0], es=rep[1], eb=rep[2], fmax=rep[3], steps=it.value)


def parse_batch(path):
    b = OrBatch()
    base = C.create_string_buffer(4096)
    tag = C.create_string_buffer(4096)
    rc = lib().or_parse_batch(os.fsencode(path), base, 4096, tag, 4096, C.byref(b))
    return rc, base.value.decode(), tag.value.decode(), b


def run_protocol(net_file, log_name, edata_name, disp_base, batch, strain_mode=True):
    return lib().or_run_protocol(os.fsencode(net_file), os.fsencode(log_name),
                                 os.fsencode(edata_name), os.fsencode(disp_base),
                                 C.byref(batch), 1 if strain_mode else 0)


def report_name(base, digits, count, ext):
    out = C.create_string_buffer(4096)
    lib().or_report_name(base.encode(), digits, count, ext.encode(), out, 4096)
    return out.value.decode()


def dec_digits(v):
    return lib().or_dec_digits(v)


def uniform(seed, counter):
    return lib().or_uniform(seed, counter)


class Mesh:
    """Owns one or_mesh: a network on the points of a gel mesh
    (src/network_mesh_relax.cu)."""

    def __init__(self):
        self.s = OrMesh()
        self._owned = False

    @classmethod
    def read(cls, path):
        m = cls()
        rc = lib().or_read_mesh(os.fsencode(path), C.byref(m.s))
        if rc:
            raise ValueError(f"or_read_mesh({path}) failed with code {rc}")
        m._owned = True
        return m

    def __del__(self):
        if getattr(self, "_owned", False) and _lib is not None:
            _lib.or_mesh_free(C.byref(self.s))
            self._owned = False

    @property
    def n(self):
        return self.s.net.n

    @property
    def e(self):
        return self.s.net.e

    @property
    def t(self):
        return self.s.net.t

    @property
    def nf(self):
        return self.s.nf

    def pos(self):
        return np.ctypeslib.as_array(self.s.net.pos, shape=(2 * self.n,)).copy()

    def bonds(self):
        return np.ctypeslib.as_array(self.s.net.bond, shape=(2 * max(self.e, 1),))[
            :2 * self.e].reshape(-1, 2).copy()

    def rest(self):
        return np.ctypeslib.as_array(self.s.net.rest, shape=(max(self.e, 1),))[:self.e].copy()

    def tris(self):
        if self.t == 0:
            return np.zeros((0, 3), np.int32)
        return np.ctypeslib.as_array(self.s.net.tri, shape=(3 * self.t,)).reshape(-1, 3).copy()

    def tri_wraps(self):
        if self.t == 0:
            return np.zeros((0, 4), np.int8)
        return np.ctypeslib.as_array(self.s.net.twrap, shape=(4 * self.t,)).reshape(-1, 4).copy()

    def fcode(self):
        return np.ctypeslib.as_array(self.s.net.fcode, shape=(self.n,)).copy()

    def facets(self):
        return np.ctypeslib.as_array(self.s.facet, shape=(3 * max(self.nf, 1),))[
            :3 * self.nf].reshape(-1, 3).copy()

    def areas(self):
        return np.ctypeslib.as_array(self.s.area, shape=(max(self.nf, 1),))[:self.nf].copy()

    def inv(self):
        return np.ctypeslib.as_array(self.s.inv, shape=(6 * max(self.nf, 1),))[
            :6 * self.nf].reshape(-1, 6).copy()

    def pairs(self):
        return np.ctypeslib.as_array(self.s.pair, shape=(2 * max(self.s.np, 1),))[
            :2 * self.s.np].reshape(-1, 2).copy()

    def setup(self, ks, kb, bmod, dilation):
        """Dilated start (2n) and preferred dilation; finds the triples (kb > 0)."""
        pos = np.zeros(2 * self.n)
        pd = C.c_double()
        rc = lib().or_mesh_setup(C.byref(self.s), ks, kb, bmod, dilation, _dp(pos),
                                 C.byref(pd))
        if rc:
            raise ValueError(f"or_mesh_setup failed with code {rc}")
        return pos, pd.value

    def inv_mats(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        lib().or_mesh_inv_mats(C.byref(self.s), _dp(pos))

    def gel_energy(self, smod, bmod, pd, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        return lib().or_gel_energy(C.byref(self.s), smod, bmod, pd, _dp(pos))

    def forces(self, ks, kb, smod, bmod, pd, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        f = np.zeros(2 * self.n)
        lib().or_mesh_forces(C.byref(self.s), ks, kb, smod, bmod, pd, _dp(pos), _dp(f))
        return f

    def energies(self, ks, kb, smod, bmod, pd, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        e = np.zeros(3)
        lib().or_mesh_energies(C.byref(self.s), ks, kb, smod, bmod, pd, _dp(pos), _dp(e))
        return e

    def affine(self, pos, xdisp, ydisp, yrange):
        lib().or_affine(C.byref(self.s.net), _dp(pos), self.s.net.miny, yrange, xdisp, ydisp)

    def fire(self, ks, kb, smod, bmod, pd, pos, fp=(0.1, 0.99, 0.1, 1.1, 0.5), nmin=5,
             mass=1.0, fcut=1e-10, max_steps=1000000):
        fpa = (C.c_double * 5)(*fp)
        rep = np.zeros(5)
        steps = C.c_int()
        rc = lib().or_mesh_fire(C.byref(self.s), ks, kb, smod, bmod, pd, _dp(pos), fpa, nmin,
                                mass, fcut, max_steps, _dp(rep), C.byref(steps))
        return dict(rc=rc, rms=rep[0], es=rep[1], eb=rep[2], eg=rep[3], fmax=rep[4],
                    steps=steps.value)


def parse_mesh_batch(path):
    b = OrMeshBatch()
    base = C.create_string_buffer(4096)
    tag = C.create_string_buffer(4096)
    rc = lib().or_parse_mesh_batch(os.fsencode(path), base, 4096, tag, 4096, C.byref(b))
    return rc, base.value.decode(), tag.value.decode(), b


def run_mesh_protocol(net_file, log_name, edata_name, disp_base, batch, strain_mode=True):
    pd = C.c_double()
    rc = lib().or_run_mesh_protocol(os.fsencode(net_file), os.fsencode(log_name),
                                    os.fsencode(edata_name), os.fsencode(disp_base),
                                    C.byref(batch), 1 if strain_mode else 0, C.byref(pd))
    return rc, pd.value
