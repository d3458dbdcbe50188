"""ctypes binding of libdgrain.so (include/dgrain.h).

The library is the product: there is no Python or CPU fallback.  Loading
fails loudly when the shared object has not been built.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)                      # .../dg-rain_amd
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libdgrain.so")
NUM_VARIANTS = 4   # DGR_NUM_VARIANTS in include/dgrain.h

DGR_OK = 0
DGR_ERR_ARG = 1
DGR_ERR_UNSUPPORTED = 7
DGR_ERR_TOO_LARGE = 8
NOALLOC_SCRATCH_BYTES = 8 << 20   # DGR_NOALLOC_SCRATCH_BYTES in include/dgrain.h
STATUS = {0: "ok", 1: "invalid argument", 2: "inconsistent mesh", 3: "HIP runtime error",
          4: "NaN flux or wet/dry failure", 5: "out of memory", 6: "no HIP device", 7: "unsupported",
          8: "request exceeds the fixed workspace"}


class DgrError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__("libdgrain: %s%s" % (STATUS.get(code, "status %d" % code), (" (%s)" % what) if what else ""))


class MeshView(C.Structure):
    _fields_ = [("num_el", C.c_int), ("num_nodes", C.c_int), ("num_edges", C.c_int),
                ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
                ("z", C.POINTER(C.c_double)), ("n", C.POINTER(C.c_double)),
                ("elto_vert", C.POINTER(C.c_int)), ("edg_to_vert", C.POINTER(C.c_int)),
                ("edg_to_els", C.POINTER(C.c_int)), ("bdry", C.POINTER(C.c_int))]


class Params(C.Structure):
    _fields_ = [("g", C.c_double), ("h0", C.c_double), ("velzero", C.c_double), ("cfl", C.c_double),
                ("dt0", C.c_double), ("final_time", C.c_double), ("rain", C.c_double),
                ("wetdry", C.c_int), ("tide_code", C.c_int),
                ("tide_datum", C.c_double), ("tide_amp", C.c_double), ("tide_period", C.c_double),
                ("tide_phase", C.c_double), ("record_modify", C.c_int), ("bc_sentinels", C.c_int)]


class AdjointSpec(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("sample_every", C.c_int), ("checkpoint_every", C.c_int),
                ("ngauges", C.c_int), ("gauge_el", C.POINTER(C.c_int)), ("obs", C.POINTER(C.c_double))]


class Report(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("max_lambda", C.c_double),
                ("error_code", C.c_int), ("error_where", C.c_int)]


_lib = None


def load():
    """Load libdgrain.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libdgrain.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
    sig = {
        "dgr_params_default": (None, [C.POINTER(Params)]),
        "dgr_status_string": (C.c_char_p, [C.c_int]),
        "dgr_build_topology": (C.c_int, [C.c_int, C.c_int, ip, dp, dp, C.c_int, ip, C.c_int, ip, C.c_int, ip,
                                         ip, ip, ip, ip, dp]),
        "dgr_create": (C.c_int, [C.POINTER(MeshView), C.POINTER(Params), C.POINTER(vp)]),
        "dgr_destroy": (C.c_int, [vp]),
        "dgr_set_state": (C.c_int, [vp, dp, dp, dp]),
        "dgr_get_state": (C.c_int, [vp, dp, dp, dp]),
        "dgr_set_wd": (C.c_int, [vp, ip]),
        "dgr_get_wd": (C.c_int, [vp, ip]),
        "dgr_get_modify": (C.c_int, [vp, C.c_int, C.POINTER(C.c_ubyte)]),
        "dgr_set_boundary": (C.c_int, [vp, dp, dp]),
        "dgr_set_time": (C.c_int, [vp, C.c_double, C.c_double]),
        "dgr_compute_rhs": (C.c_int, [vp, dp, dp, dp, dp, dp, dp, dp]),
        "dgr_limit": (C.c_int, [vp]),
        "dgr_wetdry_status": (C.c_int, [vp]),
        "dgr_pdop": (C.c_int, [vp]),
        "dgr_one_time_step": (C.c_int, [vp, C.c_double, C.c_double]),
        "dgr_run": (C.c_int, [vp, C.c_int, C.POINTER(Report)]),
        "dgr_last_error": (C.c_int, [vp, ip, ip]),
        "dgr_sync": (C.c_int, [vp]),
        "dgr_device_state": (C.c_int, [vp, C.POINTER(dp), ip]),
        "dgr_math_selftest": (C.c_int, [C.c_int, dp, dp, dp]),
        "dgr_profile_begin": (C.c_int, [vp]),
        "dgr_get_report": (C.c_int, [vp, C.POINTER(Report)]),
        "dgr_set_variant": (C.c_int, [vp, C.c_int]),
        "dgr_get_variant": (C.c_int, [vp, ip]),
        "dgr_autotune": (C.c_int, [vp, C.c_int, ip, dp]),
        "dgr_set_block": (C.c_int, [vp, C.c_int]),
        "dgr_get_block": (C.c_int, [vp, ip]),
        "dgr_autotune_block": (C.c_int, [vp, C.c_int, ip, dp]),
        "dgr_set_final_time": (C.c_int, [vp, C.c_double]),
        "dgr_couple_dgshed": (C.c_int, [vp, C.c_int, ip, ip, ip, dp, dp, dp, dp]),
        "dgr_set_graph": (C.c_int, [vp, C.c_int]),
        "dgr_ensemble_create": (C.c_int, [C.c_int, C.POINTER(MeshView), C.POINTER(Params), C.POINTER(vp)]),
        "dgr_num_members": (C.c_int, [vp, ip]),
        "dgr_member_set_state": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "dgr_member_get_state": (C.c_int, [vp, C.c_int, dp, dp, dp]),
        "dgr_member_set_wd": (C.c_int, [vp, C.c_int, ip]),
        "dgr_member_get_wd": (C.c_int, [vp, C.c_int, ip]),
        "dgr_member_get_modify": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_ubyte)]),
        "dgr_member_set_boundary": (C.c_int, [vp, C.c_int, dp, dp]),
        "dgr_member_get_boundary": (C.c_int, [vp, C.c_int, dp, dp]),
        "dgr_get_boundary": (C.c_int, [vp, dp, dp]),
        "dgr_couple_channel_ends": (C.c_int, [vp, C.c_int, ip, ip, ip, ip, dp, dp, dp, dp]),
        "dgr_set_hydrograph": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, dp]),
        "dgr_member_set_hydrograph": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp]),
        "dgr_set_infiltration": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
        "dgr_set_tape_budget": (C.c_int, [vp, C.c_longlong]),
        "dgr_member_set_infiltration": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
        "dgr_member_set_time": (C.c_int, [vp, C.c_int, C.c_double, C.c_double]),
        "dgr_member_report": (C.c_int, [vp, C.c_int, C.POINTER(Report)]),
        "dgr_member_layout": (C.c_int, [vp, C.c_int, ip, ip]),
        "dgr_adjoint_gauges": (C.c_int, [vp, C.POINTER(AdjointSpec), dp, dp, dp]),
        "dgr_gauge_run": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, ip, dp]),
        "dgr_gauge_run_noalloc": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, ip, dp]),
        "dgr_set_nodal": (C.c_int, [vp, C.c_int, dp, dp]),
        "dgr_checkpoint_write": (C.c_int, [vp, C.c_int, C.c_char_p]),
        "dgr_checkpoint_read": (C.c_int, [vp, C.c_int, C.c_char_p]),
        "dgr_profile_end": (C.c_int, [vp, dp, ip]),
        "dgr_get_stream": (C.c_int, [vp, C.POINTER(vp)]),
        "dgr_locality_order": (C.c_int, [C.c_int, C.c_int, ip, dp, dp, ip, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names declared in include/dgrain.h."""
    import re
    with open(os.path.join(REPO_ROOT, "include", "dgrain.h")) as f:
        text = f.read()
    decl = r"^\s*(?:dgr_status|void|const char \*|int|double)\s*\**\s*(dgr_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text, re.M)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def check(rc, what=""):
    if rc != DGR_OK:
        raise DgrError(rc, what)


def default_params(**kw):
    p = Params()
    load().dgr_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def locality_order(elto_vert, x, y):
    """(el_order, node_order) of dgr_locality_order: new element i is old
    element el_order[i], new node j is old node node_order[j]."""
    lib = load()
    ev = np.ascontiguousarray(elto_vert, dtype=np.int32).ravel()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    el_order = np.empty(ev.size // 3, np.int32)
    node_order = np.empty(x.size, np.int32)
    check(lib.dgr_locality_order(el_order.size, x.size, _ip(ev), _dp(x), _dp(y), _ip(el_order), _ip(node_order)),
           "dgr_locality_order")
    return el_order, node_order


def build_topology(elto_vert, x, y, in1=None, in2=None, out=None):
    """extract_mesh_attributes() in O(N) (see dgr_build_topology)."""
    lib = load()
    ev = np.ascontiguousarray(elto_vert, dtype=np.int32).ravel()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    ne, nn = ev.size // 3, x.size
    lists = [np.ascontiguousarray(a if a is not None else np.zeros(0), dtype=np.int32).ravel()
             for a in (in1, in2, out)]
    ned_max = nn + ne - 1
    e2v = np.empty(2 * ned_max, np.int32)
    e2e = np.empty(2 * ned_max, np.int32)
    bdry = np.empty(ned_max, np.int32)
    ned = C.c_int(0)
    mel = C.c_double(0)
    rc = lib.dgr_build_topology(ne, nn, _ip(ev), _dp(x), _dp(y),
                                lists[0].size // 2, _ip(lists[0]), lists[1].size // 2, _ip(lists[1]),
                                lists[2].size // 2, _ip(lists[2]),
                                _ip(e2v), _ip(e2e), _ip(bdry), C.byref(ned), C.byref(mel))
    check(rc, "dgr_build_topology")
    return e2v, e2e, bdry, mel.value


def _mesh_view(mesh, keep):
    arrs = [np.ascontiguousarray(a) for a in
            (mesh.x, mesh.y, mesh.z, mesh.n, mesh.elto_vert, mesh.edg_to_vert, mesh.edg_to_els, mesh.bdry)]
    keep.extend(arrs)
    x, y, z, n, ev, e2v, e2e, bd = arrs
    return MeshView(mesh.num_el, mesh.num_nodes, mesh.num_edges, _dp(x), _dp(y), _dp(z), _dp(n),
                    _ip(ev), _ip(e2v), _ip(e2e), _ip(bd))


class _Handle:
    """What single-instance and ensemble handles share."""

    def close(self):
        if getattr(self, "h", None):
            self.lib.dgr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps):
        """nsteps steps of every member; returns member 0's report."""
        rep = Report()
        rc = self.lib.dgr_run(self.h, int(nsteps), C.byref(rep))
        check(rc, "dgr_run (error %d at %d)" % (rep.error_code, rep.error_where))
        return rep

    def sync(self):
        check(self.lib.dgr_sync(self.h))

    def set_variant(self, v):
        check(self.lib.dgr_set_variant(self.h, int(v)))

    def variant(self):
        v = C.c_int(0)
        check(self.lib.dgr_get_variant(self.h, C.byref(v)))
        return v.value

    def set_graph(self, on=True):
        """Replay one captured hipGraph per step in run()."""
        check(self.lib.dgr_set_graph(self.h, 1 if on else 0))

    def autotune(self, steps=5):
        """Pick the fastest kernel variant; returns (chosen, ms/step of each)."""
        v = C.c_int(0)
        ms = np.zeros(NUM_VARIANTS)
        check(self.lib.dgr_autotune(self.h, int(steps), C.byref(v), _dp(ms)))
        return v.value, [float(x) for x in ms]

    def set_block(self, threads):
        """Workgroup size (64, 128, 256) of the RHS, node-bound and limiter kernels."""
        check(self.lib.dgr_set_block(self.h, int(threads)), "dgr_set_block")

    def block(self):
        t = C.c_int(0)
        check(self.lib.dgr_get_block(self.h, C.byref(t)))
        return t.value

    def autotune_block(self, steps=5):
        """Pick the fastest workgroup size for the current variant; returns
        (chosen, ms/step at 64, 128 and 256 threads)."""
        t = C.c_int(0)
        ms = np.zeros(3)
        check(self.lib.dgr_autotune_block(self.h, int(steps), C.byref(t), _dp(ms)), "dgr_autotune_block")
        return t.value, [float(x) for x in ms]

    def set_final_time(self, t):
        check(self.lib.dgr_set_final_time(self.h, float(t)))

    def profile_begin(self):
        check(self.lib.dgr_profile_begin(self.h))

    def stream(self):
        """The handle's HIP stream as an integer handle (every kernel of the
        handle runs on it), e.g. for torch.cuda.ExternalStream."""
        p = C.c_void_p()
        check(self.lib.dgr_get_stream(self.h, C.byref(p)), "dgr_get_stream")
        return p.value

    def profile_end(self):
        """(summed ms, launches) of the flux/RHS/update kernels since profile_begin."""
        ms = C.c_double(0)
        n = C.c_int(0)
        check(self.lib.dgr_profile_end(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # -- gauges and gradients ---------------------------------------------
    def _num_nodes(self):
        ms = getattr(self, "meshes", None) or [self.mesh]
        return sum(m.num_nodes for m in ms)

    def _nmem(self):
        return len(getattr(self, "meshes", None) or [self.mesh])

    def _num_edges(self, member):
        ms = getattr(self, "meshes", None) or [self.mesh]
        return ms[member].num_edges

    def gauge_run(self, nsteps, sample_every, gauge_el):
        """Run nsteps steps; gauge depths after every sample_every-th step,
        shape [members, nsteps // sample_every, ngauges]."""
        gel = np.ascontiguousarray(np.reshape(gauge_el, (self._nmem(), -1)), dtype=np.int32)
        nsamp = nsteps // sample_every
        out = np.zeros((gel.shape[0], nsamp, gel.shape[1]))
        check(self.lib.dgr_gauge_run(self.h, int(nsteps), int(sample_every), gel.shape[1], _ip(gel), _dp(out)),
              "dgr_gauge_run")
        return out

    def gauge_run_noalloc(self, nsteps, sample_every, gauge_el):
        """gauge_run through dgr_gauge_run_noalloc: no device allocation, the
        staging in the library's fixed per-device workspace."""
        gel = np.ascontiguousarray(np.reshape(gauge_el, (self._nmem(), -1)), dtype=np.int32)
        nsamp = nsteps // sample_every
        out = np.zeros((gel.shape[0], nsamp, gel.shape[1]))
        check(self.lib.dgr_gauge_run_noalloc(self.h, int(nsteps), int(sample_every), gel.shape[1], _ip(gel),
                                             _dp(out)), "dgr_gauge_run_noalloc")
        return out

    def adjoint_gauges(self, nsteps, sample_every, checkpoint_every, gauge_el, obs):
        """(loss[members], dL/dz, dL/dn) of the gauge misfit over nsteps steps
        from the current state (which is restored afterwards)."""
        gel = np.ascontiguousarray(np.reshape(gauge_el, (self._nmem(), -1)), dtype=np.int32)
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        assert obs.shape == (gel.shape[0], nsteps // sample_every, gel.shape[1]), obs.shape
        spec = AdjointSpec(int(nsteps), int(sample_every), int(checkpoint_every), gel.shape[1], _ip(gel), _dp(obs))
        loss = np.zeros(gel.shape[0])
        gz = np.zeros(self._num_nodes())
        gn = np.zeros(self._num_nodes())
        check(self.lib.dgr_adjoint_gauges(self.h, C.byref(spec), _dp(loss), _dp(gz), _dp(gn)), "dgr_adjoint_gauges")
        return loss, gz, gn

    def set_hydrograph(self, code, t, zeta_b, qn_b, member=0):
        """Boundary code `code` (1-3) follows (zeta_b, qn_b)(t), piecewise
        linear through the points; empty arrays remove it."""
        t, z, q = (np.ascontiguousarray(a, dtype=np.float64) for a in (t, zeta_b, qn_b))
        if not (t.size == z.size == q.size):
            raise ValueError("hydrograph arrays differ in length")
        check(self.lib.dgr_member_set_hydrograph(self.h, int(member), int(code), t.size, _dp(t), _dp(z), _dp(q)),
              "dgr_member_set_hydrograph")

    def get_boundary(self, member=0):
        """(bzeta, bQn) of a member's edges."""
        ned = self._num_edges(member)
        a, b = np.empty(ned), np.empty(ned)
        check(self.lib.dgr_member_get_boundary(self.h, int(member), _dp(a), _dp(b)), "dgr_member_get_boundary")
        return a, b

    def couple_channel_ends(self, end_off, edge_member, edge_local, end_dir, end_A, end_Q):
        """ChanNet's internal_BC, junction side, over the members as junctions:
        sets the coupled edges' bzeta/bQn, returns the ends' (ghost_A, ghost_Q)."""
        iv = [np.ascontiguousarray(a, np.int32) for a in (end_off, edge_member, edge_local, end_dir)]
        dv = [np.ascontiguousarray(a, np.float64) for a in (end_A, end_Q)]
        n = iv[3].size
        ga, gq = np.zeros(n), np.zeros(n)
        check(self.lib.dgr_couple_channel_ends(self.h, n, *[_ip(a) for a in iv], *[_dp(a) for a in dv], _dp(ga),
                                               _dp(gq)), "dgr_couple_channel_ends")
        return ga, gq

    def couple_dgshed(self, edge_member, edge_local, conn, chan, geom):
        """DG-SWEM's channel coupling edge, 2-D side: sets the coupled boundary
        edges' bzeta/bQn from the channel ends (chan: [n][6] = A, Q, b, m1, m2,
        z) and returns (swe [n][2] = edge-averaged zeta and len x edge-averaged
        q.n, ghost [n][2] = the ends' ghost area and discharge, from geom:
        [n][4] = b, m1, m2, z)."""
        mem = np.ascontiguousarray(edge_member, np.int32)
        loc = np.ascontiguousarray(edge_local, np.int32)
        cn = np.ascontiguousarray(conn, np.int32)
        ch = np.ascontiguousarray(chan, np.float64).reshape(-1, 6)
        gm = np.ascontiguousarray(geom, np.float64).reshape(-1, 4)
        n = mem.size
        swe, ghost = np.zeros((n, 2)), np.zeros((n, 2))
        check(self.lib.dgr_couple_dgshed(self.h, n, _ip(mem), _ip(loc), _ip(cn), _dp(ch), _dp(gm), _dp(swe),
                                         _dp(ghost)), "dgr_couple_dgshed")
        return swe, ghost

    def oneTimeStep2D(self, t, dt):
        """oneTimeStep2D with a host dt, every member."""
        check(self.lib.dgr_one_time_step(self.h, float(t), float(dt)), "dgr_one_time_step")

    def set_tape_budget(self, nbytes):
        """Device memory the adjoint may keep taped segments in (< 0: the free
        memory less 1/16 of the device's, 0: recompute every segment)."""
        check(self.lib.dgr_set_tape_budget(self.h, int(nbytes)), "dgr_set_tape_budget")

    def set_infiltration(self, K, psi=0.0, theta_e=0.0, s_e=0.0, member=0):
        """Green-Ampt infiltration of a member (K <= 0 turns it off)."""
        check(self.lib.dgr_member_set_infiltration(self.h, int(member), float(K), float(psi), float(theta_e),
                                                   float(s_e)), "dgr_member_set_infiltration")

    def checkpoint_write(self, path, member=0):
        """Restart file of a member: state, wet/dry flags, t and dt."""
        check(self.lib.dgr_checkpoint_write(self.h, int(member), os.fsencode(path)), "dgr_checkpoint_write")

    def checkpoint_read(self, path, member=0):
        check(self.lib.dgr_checkpoint_read(self.h, int(member), os.fsencode(path)), "dgr_checkpoint_read")

    def set_nodal(self, z=None, n=None, member=0):
        """Replace the nodal depth z and/or Manning n of a member."""
        if z is None and n is None:
            return
        za = None if z is None else np.ascontiguousarray(z, dtype=np.float64)
        na = None if n is None else np.ascontiguousarray(n, dtype=np.float64)
        check(self.lib.dgr_set_nodal(self.h, int(member), None if za is None else _dp(za),
                                     None if na is None else _dp(na)), "dgr_set_nodal")

    def device_state(self):
        p = C.POINTER(C.c_double)()
        pitch = C.c_int(0)
        check(self.lib.dgr_device_state(self.h, C.byref(p), C.byref(pitch)))
        return C.cast(p, C.c_void_p).value, pitch.value


class Solver(_Handle):
    """One libdgrain handle: the reference's global-state step as an object.

    Method names follow the reference functions they replace
    (compute2DL, SlopeLimiter, wetDryStatus2D, PDop2D, oneTimeStep2D)."""

    def __init__(self, mesh, params=None, **kw):
        lib = load()
        self.lib = lib
        self.mesh = mesh
        self.params = params if params is not None else default_params(**kw)
        self._keep = []
        mv = _mesh_view(mesh, self._keep)
        h = C.c_void_p()
        check(lib.dgr_create(C.byref(mv), C.byref(self.params), C.byref(h)), "dgr_create")
        self.h = h

    # -- state -------------------------------------------------------------
    def set_state(self, zeta, qx, qy):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (zeta, qx, qy)]
        check(self.lib.dgr_set_state(self.h, *[_dp(v) for v in a]), "dgr_set_state")

    def get_state(self):
        ne3 = 3 * self.mesh.num_el
        a = [np.empty(ne3) for _ in range(3)]
        check(self.lib.dgr_get_state(self.h, *[_dp(v) for v in a]), "dgr_get_state")
        return tuple(a)

    def set_wd(self, wd):
        a = np.ascontiguousarray(wd, dtype=np.int32)
        check(self.lib.dgr_set_wd(self.h, _ip(a)))

    def get_wd(self):
        a = np.empty(self.mesh.num_el, np.int32)
        check(self.lib.dgr_get_wd(self.h, _ip(a)))
        return a

    def get_modify(self, stage):
        a = np.empty(self.mesh.num_el, np.uint8)
        check(self.lib.dgr_get_modify(self.h, stage, a.ctypes.data_as(C.POINTER(C.c_ubyte))))
        return a

    def set_boundary(self, bzeta, bqn):
        a = np.ascontiguousarray(bzeta, dtype=np.float64)
        b = np.ascontiguousarray(bqn, dtype=np.float64)
        check(self.lib.dgr_set_boundary(self.h, _dp(a), _dp(b)))

    def set_time(self, t, dt):
        check(self.lib.dgr_set_time(self.h, float(t), float(dt)))

    # -- operators ---------------------------------------------------------
    def compute2DL(self):
        """RHS, edge fluxes and max_lambda of the current state (compute2DL.c:12-413)."""
        ne3 = 3 * self.mesh.num_el
        r = [np.empty(ne3) for _ in range(6)]
        ml = C.c_double(0)
        rc = self.lib.dgr_compute_rhs(self.h, *[_dp(v) for v in r], C.byref(ml))
        check(rc, "dgr_compute_rhs")
        return r[:3], r[3:], ml.value

    def SlopeLimiter(self):
        check(self.lib.dgr_limit(self.h), "dgr_limit")

    def wetDryStatus2D(self):
        check(self.lib.dgr_wetdry_status(self.h), "dgr_wetdry_status")

    def PDop2D(self):
        check(self.lib.dgr_pdop(self.h), "dgr_pdop")



class Ensemble(_Handle):
    """n independent members (own mesh, state, forcing, time and dt) advanced
    together, one launch per kernel for all of them (dgr_ensemble_create)."""

    def __init__(self, meshes, params):
        if len(meshes) != len(params) or not meshes:
            raise ValueError("one params object per mesh")
        lib = load()
        self.lib = lib
        self.meshes = list(meshes)
        self._keep = []
        mvs = (MeshView * len(meshes))(*[_mesh_view(m, self._keep) for m in meshes])
        prms = (Params * len(params))(*params)
        self.params = list(params)
        h = C.c_void_p()
        check(lib.dgr_ensemble_create(len(meshes), mvs, prms, C.byref(h)), "dgr_ensemble_create")
        self.h = h

    def __len__(self):
        return len(self.meshes)

    def set_state(self, i, zeta, qx, qy):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (zeta, qx, qy)]
        check(self.lib.dgr_member_set_state(self.h, int(i), *[_dp(v) for v in a]), "dgr_member_set_state")

    def get_state(self, i):
        ne3 = 3 * self.meshes[i].num_el
        a = [np.empty(ne3) for _ in range(3)]
        check(self.lib.dgr_member_get_state(self.h, int(i), *[_dp(v) for v in a]), "dgr_member_get_state")
        return tuple(a)

    def get_wd(self, i):
        a = np.empty(self.meshes[i].num_el, np.int32)
        check(self.lib.dgr_member_get_wd(self.h, int(i), _ip(a)))
        return a

    def get_modify(self, i, stage):
        a = np.empty(self.meshes[i].num_el, np.uint8)
        check(self.lib.dgr_member_get_modify(self.h, int(i), stage, a.ctypes.data_as(C.POINTER(C.c_ubyte))))
        return a

    def set_boundary(self, i, bzeta, bqn):
        a = np.ascontiguousarray(bzeta, dtype=np.float64)
        b = np.ascontiguousarray(bqn, dtype=np.float64)
        check(self.lib.dgr_member_set_boundary(self.h, int(i), _dp(a), _dp(b)))

    def report(self, i):
        """member i's report; raises DgrError if it recorded an error"""
        rep = Report()
        rc = self.lib.dgr_member_report(self.h, int(i), C.byref(rep))
        check(rc, "member %d (error %d at %d)" % (i, rep.error_code, rep.error_where))
        return rep

    def run(self, nsteps):
        """nsteps steps of every member; returns the per-member reports."""
        rc = self.lib.dgr_run(self.h, int(nsteps), None)
        reps = []
        for i in range(len(self)):
            rep = Report()
            self.lib.dgr_member_report(self.h, i, C.byref(rep))
            reps.append(rep)
        if rc != DGR_OK:
            bad = [(i, r.error_code, r.error_where) for i, r in enumerate(reps) if r.error_code]
            raise DgrError(rc, "dgr_run: member errors (member, code, where) %s" % bad)
        return reps


def math_selftest(x, y):
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty((x.size, 9))
    check(lib.dgr_math_selftest(x.size, _dp(x), _dp(y), _dp(out)), "dgr_math_selftest")
    return out


INF = math.inf
