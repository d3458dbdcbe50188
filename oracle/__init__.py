"""Python bindings of the CPU oracle and of the compiled reference operators.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.

  Oracle    liboracle.so (dgswe_oracle.c), the C restatement of the 2DCode step.
  Ref2D     oracle/_ref/libref2dcode.so: the reference's own 2DCode operators
            (global state, MeshAttributes.h), compiled from the reference sources.
  RefChan   oracle/_ref/librefchan2d.so: ChanNet's 2D operators built -DWDON.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the same source built at the reference's flags (-g, no optimisation)
LIB_G = os.path.join(HERE, "liboracle_g.so")
REF2D = os.path.join(HERE, "_ref", "libref2dcode.so")
REFCHAN = os.path.join(HERE, "_ref", "librefchan2d.so")

MATH_REF, MATH_DEV = 0, 1


def build(ref=True):
    """make -C oracle (the reference libraries only where the sources exist)."""
    subprocess.run(["make", "-C", HERE, "oracle"] + (["ref"] if ref else []), check=True,
                   stdout=subprocess.DEVNULL)


def ref_available():
    return os.path.exists(REF2D) and os.path.exists(REFCHAN)


class _Mesh(C.Structure):
    _fields_ = [("num_el", C.c_int), ("num_nodes", C.c_int), ("num_edges", C.c_int),
                ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
                ("z", C.POINTER(C.c_double)), ("n", C.POINTER(C.c_double)),
                ("elto_vert", C.POINTER(C.c_int)), ("edg_to_vert", C.POINTER(C.c_int)),
                ("edg_to_els", C.POINTER(C.c_int)), ("bdry", C.POINTER(C.c_int))]


class _Params(C.Structure):
    _fields_ = [("g", C.c_double), ("h0", C.c_double), ("velzero", C.c_double),
                ("wetdry", C.c_int), ("math", C.c_int),
                ("cfl", C.c_double), ("dt0", C.c_double), ("final_time", C.c_double), ("rain", C.c_double),
                ("tide_code", C.c_int), ("tide_datum", C.c_double), ("tide_amp", C.c_double),
                ("tide_period", C.c_double), ("tide_phase", C.c_double), ("no_sentinels", C.c_int)]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


_olibs = {}


def _load_oracle(path=LIB):
    if path not in _olibs:
        if not os.path.exists(path):
            build(ref=False)
        lib = C.CDLL(path)
        dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
        ub = C.POINTER(C.c_ubyte)
        sig = {
            "orc_create": (vp, [C.POINTER(_Mesh), C.POINTER(_Params)]),
            "orc_destroy": (None, [vp]),
            "orc_set_threads": (None, [vp, C.c_int]),
            "orc_set_state": (None, [vp, dp, dp, dp]),
            "orc_get_state": (None, [vp, dp, dp, dp]),
            "orc_set_wd": (None, [vp, ip]),
            "orc_get_wd": (None, [vp, ip]),
            "orc_get_modify": (None, [vp, C.c_int, ub]),
            "orc_set_boundary": (None, [vp, dp, dp]),
            "orc_get_boundary": (None, [vp, dp, dp]),
            "orc_get_fhat": (None, [vp, dp, dp, dp]),
            "orc_get_rhs": (None, [vp, dp, dp, dp]),
            "orc_max_lambda": (C.c_double, [vp]),
            "orc_min_edge_length": (C.c_double, [vp]),
            "orc_error": (C.c_int, [vp, ip]),
            "orc_roe_flux": (C.c_double, [dp] + [C.c_double] * 11),
            "orc_inner_products": (None, [vp, C.c_int, dp, dp]),
            "orc_compute_L": (C.c_int, [vp]),
            "orc_slope_limiter": (None, [vp, ub]),
            "orc_wetdry_status": (None, [vp]),
            "orc_pdop": (None, [vp]),
            "orc_pdop_changed": (C.c_longlong, [vp]),
            "orc_boundary_conditions": (None, [vp, C.c_double]),
            "orc_set_hydrograph": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, dp]),
            "orc_hydro_eval": (C.c_double, [C.c_int, dp, dp, C.c_double]),
            "orc_set_infiltration": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_ga_cumulative": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "orc_log": (C.c_double, [C.c_double]),
            "orc_log1p": (C.c_double, [C.c_double]),
            "orc_couple_ends": (C.c_int, [C.POINTER(vp), C.c_int, ip, ip, ip, ip, dp, dp, dp, dp]),
            "orc_couple_dgshed": (C.c_int, [vp, C.c_int, ip, ip, dp, dp, dp, dp]),
            "orc_one_time_step": (C.c_int, [vp, C.c_double, C.c_double]),
            "orc_run": (C.c_int, [vp, C.c_int, dp, dp]),
            "orc_pow43": (C.c_double, [C.c_double]),
            "orc_sin": (C.c_double, [C.c_double]),
            "orc_build_topology": (C.c_int, [C.c_int, C.c_int, ip, dp, dp, C.c_int, ip, C.c_int, ip, C.c_int, ip,
                                             ip, ip, ip, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _olibs[path] = lib
    return _olibs[path]


def hydro_eval(t, v, time):
    t = np.ascontiguousarray(t, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    return _load_oracle().orc_hydro_eval(t.size, _dp(t), _dp(v), float(time))


def ga_cumulative(K, P, t):
    return _load_oracle().orc_ga_cumulative(float(K), float(P), float(t))


def dlog(x):
    return _load_oracle().orc_log(float(x))


def dlog1p(x):
    return _load_oracle().orc_log1p(float(x))


def couple_dgshed(o, edges, conn, chan, geom):
    """DG-SWEM's channel coupling edge, 2-D side, on an Oracle: sets the
    coupled boundary edges' bzeta/bQn from the channel ends (chan: [n][6] =
    A, Q, b, m1, m2, z) and returns (swe [n][2] = edge-averaged zeta and len x
    edge-averaged q.n, ghost [n][2] = the ends' ghost area and discharge,
    from geom: [n][4] = b, m1, m2, z)."""
    lib = _load_oracle()
    e = np.ascontiguousarray(edges, np.int32)
    c = np.ascontiguousarray(conn, np.int32)
    ch = np.ascontiguousarray(chan, np.float64).reshape(-1, 6)
    gm = np.ascontiguousarray(geom, np.float64).reshape(-1, 4)
    n = e.size
    swe, ghost = np.zeros((n, 2)), np.zeros((n, 2))
    if lib.orc_couple_dgshed(o.h, n, _ip(e), _ip(c), _dp(ch), _dp(gm), _dp(swe), _dp(ghost)):
        raise ValueError("a coupled edge is not a boundary edge")
    return swe, ghost


def couple_ends(juncs, end_off, edge_junc, edge_local, end_dir, end_A, end_Q):
    """ChanNet's internal_BC, junction side, over Oracle objects (one per
    junction): sets their bzeta/bQn at the coupled edges, returns the ends'
    (ghost_A, ghost_Q)."""
    lib = _load_oracle()
    iv = [np.ascontiguousarray(a, np.int32) for a in (end_off, edge_junc, edge_local, end_dir)]
    dv = [np.ascontiguousarray(a, np.float64) for a in (end_A, end_Q)]
    n = iv[3].size
    ga, gq = np.zeros(n), np.zeros(n)
    hs = (C.c_void_p * len(juncs))(*[o.h for o in juncs])
    rc = lib.orc_couple_ends(hs, n, *[_ip(a) for a in iv], *[_dp(a) for a in dv], _dp(ga), _dp(gq))
    if rc:
        raise ValueError("a coupled edge is not an edge of its element")
    return ga, gq


def pow43(x):
    return _load_oracle().orc_pow43(float(x))


def dsin(x):
    return _load_oracle().orc_sin(float(x))


def roe_flux(zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, lg, g=9.81):
    F = np.zeros(3)
    lam = _load_oracle().orc_roe_flux(_dp(F), zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, lg, g)
    return F, lam


def build_topology(elto_vert, x, y, in1=(), in2=(), out=()):
    lib = _load_oracle()
    ev = np.ascontiguousarray(elto_vert, np.int32).ravel()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    ne, nn = ev.size // 3, x.size
    L = [np.ascontiguousarray(np.asarray(a).ravel(), np.int32) for a in (in1, in2, out)]
    m = nn + ne - 1
    e2v, e2e, bd = np.empty(2 * m, np.int32), np.empty(2 * m, np.int32), np.empty(m, np.int32)
    mel = C.c_double()
    ned = lib.orc_build_topology(ne, nn, _ip(ev), _dp(x), _dp(y), L[0].size // 2, _ip(L[0]),
                                 L[1].size // 2, _ip(L[1]), L[2].size // 2, _ip(L[2]),
                                 _ip(e2v), _ip(e2e), _ip(bd), C.byref(mel))
    if ned < 0:
        raise ValueError("orc_build_topology failed (%d)" % ned)
    return e2v, e2e, bd, mel.value


class Oracle:
    """The C restatement as an object (mirrors dgrain.native.Solver)."""

    def __init__(self, mesh, math=MATH_DEV, wetdry=0, cfl=0.1, dt0=1e-7, final_time=np.inf, rain=0.0,
                 tide_code=0, tide_datum=0.0, tide_amp=0.0, tide_period=43200.0, tide_phase=0.0,
                 g=9.81, h0=1e-3, velzero=1e-2, threads=1, bc_sentinels=1, lib=LIB, **_ignored):
        lib = _load_oracle(lib)
        self.lib = lib
        self.mesh = mesh
        self._keep = [np.ascontiguousarray(a) for a in
                      (mesh.x, mesh.y, mesh.z, mesh.n, mesh.elto_vert, mesh.edg_to_vert, mesh.edg_to_els, mesh.bdry)]
        x, y, z, n, ev, e2v, e2e, bd = self._keep
        m = _Mesh(mesh.num_el, mesh.num_nodes, mesh.num_edges, _dp(x), _dp(y), _dp(z), _dp(n),
                  _ip(ev), _ip(e2v), _ip(e2e), _ip(bd))
        p = _Params(g, h0, velzero, wetdry, math, cfl, dt0, final_time, rain, tide_code, tide_datum,
                    tide_amp, tide_period, tide_phase, 0 if bc_sentinels else 1)
        self.p = p
        self.h = lib.orc_create(C.byref(m), C.byref(p))
        if threads > 1:
            lib.orc_set_threads(self.h, threads)
        self.t = 0.0
        self.dt = dt0

    @classmethod
    def for_workload(cls, wl, math=MATH_DEV, threads=1, lib=LIB):
        o = cls(wl.mesh, math=math, threads=threads, lib=lib, **wl.params)
        o.set_state(wl.zeta, wl.qx, wl.qy)
        return o

    def close(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, zeta, qx, qy):
        a = [np.ascontiguousarray(v, np.float64) for v in (zeta, qx, qy)]
        self.lib.orc_set_state(self.h, *[_dp(v) for v in a])

    def get_state(self):
        a = [np.empty(3 * self.mesh.num_el) for _ in range(3)]
        self.lib.orc_get_state(self.h, *[_dp(v) for v in a])
        return tuple(a)

    def set_wd(self, wd):
        a = np.ascontiguousarray(wd, np.int32)
        self.lib.orc_set_wd(self.h, _ip(a))

    def get_wd(self):
        a = np.empty(self.mesh.num_el, np.int32)
        self.lib.orc_get_wd(self.h, _ip(a))
        return a

    def get_modify(self, stage):
        a = np.empty(self.mesh.num_el, np.uint8)
        self.lib.orc_get_modify(self.h, stage, a.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return a

    def set_boundary(self, bz, bq):
        a, b = np.ascontiguousarray(bz, np.float64), np.ascontiguousarray(bq, np.float64)
        self.lib.orc_set_boundary(self.h, _dp(a), _dp(b))

    def get_boundary(self):
        a, b = np.empty(self.mesh.num_edges), np.empty(self.mesh.num_edges)
        self.lib.orc_get_boundary(self.h, _dp(a), _dp(b))
        return a, b

    def boundary_conditions(self, t):
        self.lib.orc_boundary_conditions(self.h, float(t))

    def set_infiltration(self, K, psi, theta_e, s_e=0.0):
        if self.lib.orc_set_infiltration(self.h, float(K), float(psi), float(theta_e), float(s_e)):
            raise ValueError("bad infiltration parameters")

    def set_hydrograph(self, code, t, zeta, qn):
        t, z, q = (np.ascontiguousarray(a, np.float64) for a in (t, zeta, qn))
        if not (t.size == z.size == q.size):
            raise ValueError("hydrograph arrays differ in length")
        if self.lib.orc_set_hydrograph(self.h, int(code), t.size, _dp(t), _dp(z), _dp(q)):
            raise ValueError("bad hydrograph")

    def compute2DL(self):
        rc = self.lib.orc_compute_L(self.h)
        if rc:
            raise RuntimeError("oracle compute2DL error %d" % rc)
        ne3 = 3 * self.mesh.num_el
        r = [np.empty(ne3) for _ in range(6)]
        self.lib.orc_get_rhs(self.h, *[_dp(v) for v in r[:3]])
        self.lib.orc_get_fhat(self.h, *[_dp(v) for v in r[3:]])
        return r[:3], r[3:], self.lib.orc_max_lambda(self.h)

    def inner_products(self, el):
        SI, VI = np.empty(9), np.empty(21)
        self.lib.orc_inner_products(self.h, el, _dp(SI), _dp(VI))
        return SI, VI

    def SlopeLimiter(self):
        mod = np.empty(self.mesh.num_el, np.uint8)
        self.lib.orc_slope_limiter(self.h, mod.ctypes.data_as(C.POINTER(C.c_ubyte)))
        return mod

    def wetDryStatus2D(self):
        self.lib.orc_wetdry_status(self.h)

    def PDop2D(self):
        self.lib.orc_pdop(self.h)
        w = C.c_int()
        rc = self.lib.orc_error(self.h, C.byref(w))
        if rc:
            raise RuntimeError("oracle PDop error %d at %d" % (rc, w.value))

    def oneTimeStep2D(self, t, dt):
        rc = self.lib.orc_one_time_step(self.h, float(t), float(dt))
        if rc:
            raise RuntimeError("oracle step error %d" % rc)

    def run(self, nsteps):
        t, dt = C.c_double(self.t), C.c_double(self.dt)
        rc = self.lib.orc_run(self.h, int(nsteps), C.byref(t), C.byref(dt))
        self.t, self.dt = t.value, dt.value
        if rc:
            w = C.c_int()
            self.lib.orc_error(self.h, C.byref(w))
            raise RuntimeError("oracle run error %d at %d" % (rc, w.value))
        return self.t, self.dt

    @property
    def max_lambda(self):
        return self.lib.orc_max_lambda(self.h)

    @property
    def pdop_changed(self):
        """Elements whose state PDop2D changed, summed over all its calls."""
        return self.lib.orc_pdop_changed(self.h)

    @property
    def min_edge_length(self):
        return self.lib.orc_min_edge_length(self.h)


# ---------------------------------------------------------------------------
# compiled reference operators
# ---------------------------------------------------------------------------
class Ref2D:
    """The reference's 2DCode operators over its globals (MeshAttributes.h:5-42).

    One instance at a time (the reference is global state).  The RK
    composition of oneTimeStep2D (time_evolution.c:10-246, which needs GSL) is
    restated here with numpy's exact elementwise IEEE operations."""

    _lib = None

    def __init__(self, mesh, wetdry=False):
        if Ref2D._lib is None:
            if not os.path.exists(REF2D):
                raise RuntimeError("oracle/_ref/libref2dcode.so not built")
            Ref2D._lib = C.CDLL(REF2D)
            lib = Ref2D._lib
            lib.RoeFlux.restype = C.c_double
            lib.RoeFlux.argtypes = [C.POINTER(C.c_double)] + [C.c_double] * 10
            lib.Compute2DInnerProducts.restype = None
            lib.Compute2DInnerProducts.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_double]
            lib.compute2DL.argtypes = [C.c_double]
            lib.store_mesh.argtypes = [C.c_char_p]
        self.lib = lib = Ref2D._lib
        self.mesh = mesh
        self.wetdry = wetdry
        ne, nn, ned = mesh.num_el, mesh.num_nodes, mesh.num_edges
        self.arrs = {
            "x_coord": np.ascontiguousarray(mesh.x, np.float64).copy(),
            "y_coord": np.ascontiguousarray(mesh.y, np.float64).copy(),
            "z": np.ascontiguousarray(mesh.z, np.float64).copy(),
            "nFriction": np.ascontiguousarray(mesh.n, np.float64).copy(),
            "EltoVert": np.ascontiguousarray(mesh.elto_vert, np.int32).copy(),
            "EdgtoEls": np.ascontiguousarray(mesh.edg_to_els, np.int32).copy(),
            "EdgtoVert": np.ascontiguousarray(mesh.edg_to_vert, np.int32).copy(),
            "BdryPrescribed": np.ascontiguousarray(mesh.bdry, np.int32).copy(),
            "zeta": np.zeros(3 * ne), "Qx": np.zeros(3 * ne), "Qy": np.zeros(3 * ne),
            "Fhat1dotn": np.zeros(3 * ne), "Fhat2dotn": np.zeros(3 * ne), "Fhat3dotn": np.zeros(3 * ne),
            "RHSZeta": np.zeros(3 * ne), "RHSQx": np.zeros(3 * ne), "RHSQy": np.zeros(3 * ne),
            "bzeta": np.zeros(ned), "bQn": np.zeros(ned),
            "WD": np.full(ne, -1, np.int32),
        }
        for name, a in self.arrs.items():
            C.c_void_p.in_dll(lib, name).value = a.ctypes.data
        C.c_int.in_dll(lib, "NumEl").value = ne
        C.c_int.in_dll(lib, "NumNodes").value = nn
        C.c_int.in_dll(lib, "NumEdges").value = ned
        C.c_double.in_dll(lib, "minEdgLength").value = mesh.min_edge_length
        lib.boundary_conditions()

    def set_state(self, zeta, qx, qy):
        self.arrs["zeta"][:] = zeta
        self.arrs["Qx"][:] = qx
        self.arrs["Qy"][:] = qy

    def get_state(self):
        return self.arrs["zeta"].copy(), self.arrs["Qx"].copy(), self.arrs["Qy"].copy()

    def set_wd(self, wd):
        self.arrs["WD"][:] = wd

    def get_wd(self):
        return self.arrs["WD"].copy()

    def set_boundary(self, bz, bq):
        self.arrs["bzeta"][:] = bz
        self.arrs["bQn"][:] = bq

    def get_boundary(self):
        return self.arrs["bzeta"].copy(), self.arrs["bQn"].copy()

    @property
    def max_lambda(self):
        return C.c_double.in_dll(self.lib, "max_lambda").value

    def compute2DL(self):
        self.lib.compute2DL(0.0)
        a = self.arrs
        return ([a["RHSZeta"].copy(), a["RHSQx"].copy(), a["RHSQy"].copy()],
                [a["Fhat1dotn"].copy(), a["Fhat2dotn"].copy(), a["Fhat3dotn"].copy()], self.max_lambda)

    def inner_products(self, el):
        SI, VI = np.zeros(9), np.zeros(21)
        self.lib.Compute2DInnerProducts(_dp(SI), _dp(VI), int(el), 0.0)
        return SI, VI

    def SlopeLimiter(self):
        self.lib.SlopeLimiter()

    def wetDryStatus2D(self):
        self.lib.wetDryStatus2D()

    def PDop2D(self):
        self.lib.PDop2D()

    def oneTimeStep2D(self, t, dt):
        """time_evolution.c:10-246 over the reference operators (no wet/dry:
        2DCode's calls are commented out, :41,149,181,233)."""
        a = self.arrs
        w = [a["zeta"].copy(), a["Qx"].copy(), a["Qy"].copy()]
        self.lib.compute2DL(t)
        for f, (sk, rk) in enumerate((("zeta", "RHSZeta"), ("Qx", "RHSQx"), ("Qy", "RHSQy"))):
            a[sk][:] = w[f] + a[rk] * dt
        self.lib.SlopeLimiter()
        w1 = [a["zeta"] + w[0], a["Qx"] + w[1], a["Qy"] + w[2]]
        self.lib.compute2DL(t)
        for f, (sk, rk) in enumerate((("zeta", "RHSZeta"), ("Qx", "RHSQx"), ("Qy", "RHSQy"))):
            a[sk][:] = w1[f] * 0.5 + a[rk] * (dt / 2)
        self.lib.SlopeLimiter()

    def topology_from_fort14(self, path):
        """store_mesh + extract_mesh_attributes on a file; returns the globals."""
        lib = self.lib
        lib.store_mesh(path.encode())
        C.c_double.in_dll(lib, "minEdgLength").value = 9999999
        lib.extract_mesh_attributes()
        ned = C.c_int.in_dll(lib, "NumEdges").value
        def arr(name, n, ct):
            p = C.cast(C.c_void_p.in_dll(lib, name).value, C.POINTER(ct))
            return np.ctypeslib.as_array(p, shape=(n,)).copy()
        out = dict(num_edges=ned,
                   edg_to_vert=arr("EdgtoVert", 2 * ned, C.c_int),
                   edg_to_els=arr("EdgtoEls", 2 * ned, C.c_int),
                   bdry=arr("BdryPrescribed", ned, C.c_int),
                   min_edge_length=C.c_double.in_dll(lib, "minEdgLength").value)
        # restore this object's globals
        for name, a in self.arrs.items():
            C.c_void_p.in_dll(lib, name).value = a.ctypes.data
        C.c_int.in_dll(lib, "NumEl").value = self.mesh.num_el
        C.c_int.in_dll(lib, "NumNodes").value = self.mesh.num_nodes
        C.c_int.in_dll(lib, "NumEdges").value = self.mesh.num_edges
        C.c_double.in_dll(lib, "minEdgLength").value = self.mesh.min_edge_length
        return out


def ref_roe_flux(zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, lg):
    if Ref2D._lib is None:
        raise RuntimeError("construct a Ref2D first")
    F = np.zeros(3)
    lam = Ref2D._lib.RoeFlux(_dp(F), zin, zex, qxin, qxex, qyin, qyex, zb, nx, ny, lg)
    return F, lam


class RefChan:
    """ChanNet's 2D operators built -DWDON, driven through one struct junction."""

    _lib = None

    def __init__(self, mesh):
        if RefChan._lib is None:
            if not os.path.exists(REFCHAN):
                raise RuntimeError("oracle/_ref/librefchan2d.so not built")
            lib = C.CDLL(REFCHAN)
            dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
            lib.refchan_create.restype = vp
            lib.refchan_create.argtypes = [C.c_int] * 3 + [dp] * 4 + [ip] * 4 + [dp] * 5 + [ip]
            lib.refchan_destroy.argtypes = [vp]
            lib.refchan_compute_L.restype = C.c_double
            lib.refchan_compute_L.argtypes = [vp, dp, dp, dp]
            for nm in ("refchan_limiter", "refchan_wd_status", "refchan_pdop"):
                getattr(lib, nm).argtypes = [vp]
            lib.refchan_channel.restype = vp
            lib.refchan_channel.argtypes = [C.c_int, dp, dp, dp, C.c_int, ip, C.c_int, ip]
            lib.refchan_channel_destroy.argtypes = [vp]
            lib.refchan_internal_bc.argtypes = [C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp)]
            RefChan._lib = lib
        self.lib = lib = RefChan._lib
        self.mesh = mesh
        ne, ned = mesh.num_el, mesh.num_edges
        self.a = dict(x=np.array(mesh.x, np.float64), y=np.array(mesh.y, np.float64),
                      z=np.array(mesh.z, np.float64), n=np.array(mesh.n, np.float64),
                      ev=np.array(mesh.elto_vert, np.int32), e2v=np.array(mesh.edg_to_vert, np.int32),
                      e2e=np.array(mesh.edg_to_els, np.int32), bd=np.array(mesh.bdry, np.int32),
                      zeta=np.zeros(3 * ne), qx=np.zeros(3 * ne), qy=np.zeros(3 * ne),
                      bz=np.zeros(ned), bq=np.zeros(ned), wd=np.full(ne, -1, np.int32),
                      rz=np.zeros(3 * ne), rqx=np.zeros(3 * ne), rqy=np.zeros(3 * ne))
        a = self.a
        self.j = lib.refchan_create(ne, mesh.num_nodes, ned, _dp(a["x"]), _dp(a["y"]), _dp(a["z"]), _dp(a["n"]),
                                    _ip(a["ev"]), _ip(a["e2v"]), _ip(a["e2e"]), _ip(a["bd"]),
                                    _dp(a["zeta"]), _dp(a["qx"]), _dp(a["qy"]), _dp(a["bz"]), _dp(a["bq"]),
                                    _ip(a["wd"]))

    def __del__(self):
        try:
            self.lib.refchan_destroy(self.j)
        except Exception:
            pass

    def set_state(self, zeta, qx, qy):
        self.a["zeta"][:] = zeta
        self.a["qx"][:] = qx
        self.a["qy"][:] = qy

    def get_state(self):
        return self.a["zeta"].copy(), self.a["qx"].copy(), self.a["qy"].copy()

    def set_wd(self, wd):
        self.a["wd"][:] = wd

    def get_wd(self):
        return self.a["wd"].copy()

    def compute2DL(self):
        a = self.a
        ml = self.lib.refchan_compute_L(self.j, _dp(a["rz"]), _dp(a["rqx"]), _dp(a["rqy"]))
        return [a["rz"].copy(), a["rqx"].copy(), a["rqy"].copy()], ml

    def SlopeLimiter(self):
        self.lib.refchan_limiter(self.j)

    def get_boundary(self):
        return self.a["bz"].copy(), self.a["bq"].copy()

    @staticmethod
    def internal_bc(juncs, end_off, edge_junc, edge_local, end_dir, end_A, end_Q):
        """The reference's internal_BC over RefChan junctions, one two-node
        channel per end (its A[1], Q[1] or A[2], Q[2] the end values): sets
        the junctions' bzeta/bQn, returns the ends' (ghost_A, ghost_Q)."""
        lib = RefChan._lib
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        keep, chans = [], []
        for e in range(len(end_dir)):
            j0, j1 = int(end_off[e]), int(end_off[e + 1])
            lst = np.array([[edge_local[j], edge_junc[j]] for j in range(j0, j1)], np.int32).ravel()
            A, Q, b = np.zeros(4), np.zeros(4), np.ones(2)
            at = 1 if end_dir[e] > 0 else 2
            A[at], Q[at] = end_A[e], end_Q[e]
            keep += [lst, A, Q, b]
            n = j1 - j0
            if end_dir[e] > 0:
                c = lib.refchan_channel(2, _dp(A), _dp(Q), _dp(b), n, lst.ctypes.data_as(ip), 0, None)
            else:
                c = lib.refchan_channel(2, _dp(A), _dp(Q), _dp(b), 0, None, n, lst.ctypes.data_as(ip))
            chans.append(c)
        cs = (C.c_void_p * len(chans))(*chans)
        js = (C.c_void_p * len(juncs))(*[j.j for j in juncs])
        lib.refchan_internal_bc(len(chans), cs, len(juncs), js)
        ga = np.array([keep[4 * e + 1][0 if end_dir[e] > 0 else 3] for e in range(len(end_dir))])
        gq = np.array([keep[4 * e + 2][0 if end_dir[e] > 0 else 3] for e in range(len(end_dir))])
        for c in chans:
            lib.refchan_channel_destroy(c)
        return ga, gq

    def wetDryStatus2D(self):
        self.lib.refchan_wd_status(self.j)

    def PDop2D(self):
        self.lib.refchan_pdop(self.j)

    def oneTimeStep2D(self, t, dt):
        """ChanNet/time_evolution.c:218-427 (WDON) over the reference operators."""
        a = self.a
        self.wetDryStatus2D()
        w = [a["zeta"].copy(), a["qx"].copy(), a["qy"].copy()]
        r, _ = self.compute2DL()
        for f, k in enumerate(("zeta", "qx", "qy")):
            a[k][:] = w[f] + r[f] * dt
        self.SlopeLimiter()
        self.PDop2D()
        self.wetDryStatus2D()
        w1 = [a["zeta"] + w[0], a["qx"] + w[1], a["qy"] + w[2]]
        r, ml = self.compute2DL()
        for f, k in enumerate(("zeta", "qx", "qy")):
            a[k][:] = w1[f] * 0.5 + r[f] * (dt / 2)
        self.SlopeLimiter()
        self.PDop2D()
        return ml
