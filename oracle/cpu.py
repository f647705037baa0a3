"""ORACLE — test infrastructure only, never the product.

ctypes bindings for the CPU restatement (``oracle/liboracle.so``, prefix ``orc_``) and for the
reference's own solver compiled from its sources (``oracle/_ref/libpsref.so``, prefix ``ref_``).
Both expose the same entry points, so a test can run one case through either and compare.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {"orc": os.path.join(HERE, "liboracle.so"), "ref": os.path.join(HERE, "_ref", "libpsref.so")}

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_libs = {}


def available(kind):
    return os.path.exists(PATHS[kind])


def lib(kind):
    """The loaded library for ``kind`` in {"orc", "ref"} with argtypes set."""
    if kind in _libs:
        return _libs[kind]
    if not available(kind):
        raise FileNotFoundError(PATHS[kind] + " is not built (run make -C oracle)")
    L = C.CDLL(PATHS[kind])
    p = kind + "_"
    vp, i64, dbl = C.c_void_p, C.c_int64, C.c_double
    sig = {
        "last_error": (C.c_char_p, []),
        "hex_element": (C.c_int, [dbl, _f64p, _f64p, _f64p]),
        "create": (vp, [C.c_int, C.c_int, C.c_int, _f32p, dbl, _u8p, i64, vp]),
        "destroy": (None, [vp]),
        "num_levels": (C.c_int, [vp]),
        "level_info": (None, [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
        "level_arrays": (None, [vp, C.c_int, vp, vp, vp, vp]),
        "level_kb": (None, [vp, C.c_int, vp, vp]),
        "apply": (None, [vp, C.c_int, _f64p, _f64p, C.c_int]),
        "apply32": (None, [vp, C.c_int, _f32p, _f32p, C.c_int]),
        "restrict": (None, [vp, C.c_int, _f64p, _f64p, C.c_int]),
        "prolong_add": (None, [vp, C.c_int, _f64p, _f64p]),
        "prepare_smoothers": (None, [vp]),
        "coarse_m": (i64, [vp]),
        "coarse_solve64": (None, [vp, _f64p, _f64p]),
        "precondition": (None, [vp, _f64p, _f64p]),
        "solve": (C.c_int, [vp, _f64p, vp, dbl, C.c_int, _f64p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "reactions": (None, [vp, _f64p, vp, _f64p]),
        "stresses": (None, [vp, _f64p, dbl, dbl, C.c_int, _f32p, vp, vp]),
        "stress_tensors": (None, [vp, _f64p, dbl, dbl, _f32p]),
        "principal": (None, [dbl] * 6 + [_f64p]),
        "von_mises": (dbl, [_f64p]),
        "prune_floating": (C.c_int, [C.c_int, C.c_int, C.c_int, _f32p, _u8p, _f32p]),
        "tridiagonal_max": (dbl, [_f64p, C.c_int, _f64p]),
        "cheb_coefficients": (None, [dbl, _f64p]),
        "grid_for_box": (None, [_f64p, _f64p, C.c_int, C.c_int, _f64p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "element_energies": (None, [vp, _f64p, _f64p]),
        "optimize_topology": (C.c_int, [vp, _f32p, _u8p, _u8p, _f64p, dbl, dbl, dbl, dbl, dbl, dbl, C.c_int, i64, _f32p, _f64p]),
        "voxelize": (C.c_int, [_f32p, i64, vp, i64, _f64p, dbl, C.POINTER(C.c_int), C.c_int, _f32p]),
        "break_test": (C.c_int, [vp, _f32p, dbl, _u8p, _f64p, dbl, dbl, dbl, C.c_int, C.c_int, dbl, C.c_int, i64, _f64p, vp, vp,
                                 C.c_int, vp, C.c_int, vp, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    }
    if kind == "ref":
        sig["timed_solve"] = (dbl, [C.c_int, C.c_int, C.c_int, _f32p, dbl, _u8p, i64, _f64p, dbl, C.c_int, _f64p,
                                    C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)])
        sig["threads"] = (C.c_uint, [])
        sig["timed_phases"] = (C.c_int, [C.c_int, C.c_int, C.c_int, _f32p, dbl, _u8p, i64, _f64p, dbl, C.c_int, _f64p,
                                         C.POINTER(C.c_int), C.POINTER(C.c_double), _f64p])
    else:
        sig["topo_filter"] = (None, [vp, _f32p, dbl, C.c_int, _f64p, _f64p])
        sig["set_threads"] = (None, [C.c_int])
        sig["set_cheb_degree"] = (None, [vp, C.c_int])
        sig["smooth"] = (None, [vp, C.c_int, _f32p, _f32p])
    for name, (res, args) in sig.items():
        fn = getattr(L, p + name)
        fn.restype = res
        fn.argtypes = args
    _libs[kind] = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def hex_element(nu, kind="orc"):
    L = lib(kind)
    K = np.zeros(576)
    cs = np.zeros(8 * 144)
    gb = np.zeros(8 * 144)
    if getattr(L, kind + "_hex_element")(nu, K, cs, gb) != 0:
        raise ValueError(getattr(L, kind + "_last_error")().decode())
    return K.reshape(24, 24), cs.reshape(8, 6, 24), gb.reshape(8, 6, 24)


class Model:
    """One VoxelFEA (hierarchy + coarse factor) built by the restatement or the reference."""

    def __init__(self, dims, density, nu, bc, coarsest_max_dof=1100, diag_add=None, kind="orc"):
        self.kind = kind
        self.L = lib(kind)
        self.dims = tuple(int(d) for d in dims)
        nx, ny, nz = self.dims
        self.density = np.ascontiguousarray(density, np.float32).reshape(-1)
        self.bc = np.ascontiguousarray(bc, np.uint8).reshape(-1)
        self._da = None if diag_add is None else np.ascontiguousarray(diag_add, np.float64)
        self.h = self._f("create")(nx, ny, nz, self.density, float(nu), self.bc, int(coarsest_max_dof), _ptr(self._da))
        if not self.h:
            raise ValueError(self._f("last_error")().decode())
        self.nu = nu

    def _f(self, name):
        return getattr(self.L, self.kind + "_" + name)

    def __del__(self):
        if getattr(self, "h", None):
            self._f("destroy")(self.h)
            self.h = None

    @property
    def num_levels(self):
        return self._f("num_levels")(self.h)

    def level(self, l):
        dims = (C.c_int * 6)()
        info = (C.c_int64 * 4)()
        lm = C.c_double()
        self._f("level_info")(self.h, l, dims, info, C.byref(lm))
        return {"dims": tuple(dims[:3]), "NXYZ": tuple(dims[3:]), "nNodes": info[0], "nDof": info[1],
                "freeDof": info[2], "nOwn": info[3], "lmax": lm.value}

    @property
    def n_dof(self):
        return self.level(0)["nDof"]

    def level_arrays(self, l):
        lv = self.level(l)
        fixed = np.zeros(lv["nDof"], np.uint8)
        inv = np.zeros(lv["nDof"])
        ns = np.zeros(lv["nNodes"], np.float32)
        vs = np.zeros(int(np.prod(lv["dims"])), np.float32)
        self._f("level_arrays")(self.h, l, _ptr(fixed), _ptr(inv), _ptr(ns), _ptr(vs))
        return fixed, inv, ns, vs

    def level_kb(self, l):
        kb = np.zeros(576)
        sb = np.zeros(324)
        self._f("level_kb")(self.h, l, _ptr(kb), _ptr(sb))
        return kb.reshape(24, 24), sb

    def apply(self, x, level=0, raw=False):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        self._f("apply")(self.h, level, x, y, int(raw))
        return y

    def apply32(self, x, level=0, raw=False):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros_like(x)
        self._f("apply32")(self.h, level, x, y, int(raw))
        return y

    def restrict(self, rf, level, zero_fixed=True):
        rf = np.ascontiguousarray(rf, np.float64)
        rc = np.zeros(self.level(level + 1)["nDof"])
        self._f("restrict")(self.h, level, rf, rc, int(zero_fixed))
        return rc

    def prolong_add(self, zc, zf, level):
        zf = np.array(zf, np.float64, copy=True)
        self._f("prolong_add")(self.h, level, np.ascontiguousarray(zc, np.float64), zf)
        return zf

    def set_cheb_degree(self, degree):
        self._f("set_cheb_degree")(self.h, int(degree))

    def smooth(self, r, level=0):
        """The Chebyshev smoother's pre-smoothing pass from z = 0 (FP32)."""
        r = np.ascontiguousarray(r, np.float32)
        z = np.zeros_like(r)
        self._f("smooth")(self.h, int(level), r, z)
        return z

    def prepare_smoothers(self):
        self._f("prepare_smoothers")(self.h)
        return [self.level(l)["lmax"] for l in range(self.num_levels)]

    @property
    def coarse_m(self):
        return self._f("coarse_m")(self.h)

    def coarse_solve64(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        self._f("coarse_solve64")(self.h, r, z)
        return z

    def precondition(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        self._f("precondition")(self.h, r, z)
        return z

    def solve(self, f, tol=1e-6, max_iter=500, x0=None):
        f = np.ascontiguousarray(f, np.float64)
        x0a = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        u = np.zeros_like(f)
        its = C.c_int()
        res = C.c_double()
        ok = self._f("solve")(self.h, f, _ptr(x0a), float(tol), int(max_iter), u, C.byref(its), C.byref(res))
        if ok < 0:
            raise ValueError(self._f("last_error")().decode())
        return {"u": u, "iterations": its.value, "residual": res.value, "converged": ok == 1}

    def element_energies(self, u):
        """u_e^T K0 u_e per finest-level element (voxel order)."""
        out = np.zeros(int(np.count_nonzero(self.density > 0)))
        self._f("element_energies")(self.h, np.ascontiguousarray(u, np.float64), out)
        return out

    def reactions(self, u, f=None):
        R = np.zeros(3)
        fa = None if f is None else np.ascontiguousarray(f, np.float64)
        self._f("reactions")(self.h, np.ascontiguousarray(u, np.float64), _ptr(fa), R)
        return R

    def stresses(self, u, E, h, principals=True):
        nN = self.level(0)["nNodes"]
        vm = np.zeros(nN, np.float32)
        p1 = np.zeros(nN, np.float32) if principals else None
        p3 = np.zeros(nN, np.float32) if principals else None
        self._f("stresses")(self.h, np.ascontiguousarray(u, np.float64), float(E), float(h), int(principals), vm, _ptr(p1), _ptr(p3))
        return vm, p1, p3

    def stress_tensors(self, u, E, h):
        out = np.zeros(6 * self.level(0)["nNodes"], np.float32)
        self._f("stress_tensors")(self.h, np.ascontiguousarray(u, np.float64), float(E), float(h), out)
        return out.reshape(-1, 6)


def prune_floating(dims, density, held, kind="orc"):
    L = lib(kind)
    density = np.ascontiguousarray(density, np.float32).reshape(-1)
    out = np.zeros_like(density)
    removed = getattr(L, kind + "_prune_floating")(*[int(d) for d in dims], density, np.ascontiguousarray(held, np.uint8), out)
    return out, removed


BREAK_REASONS = ("step limit", "separated", "unsupported load", "solver did not converge", "no stress")


def break_test(dims, density, nu, bc, f, E, h, strength, principal=False, max_steps=5, tol=1e-6, max_iter=400, coarsest_max_dof=1100,
               fields=False, kind="orc"):
    """breakTest (static_study.cpp:197-279) with the VoxelFEA solve: dict(reason, peak, removed, steps), each
    step dict(lambda, max_disp, voxels, cracked, detached[, vm]). kind "ref" runs the reference's own
    breakTest without a GPU, which fixes tol 1e-6, max_iter 400 and coarsest_max_dof 1100."""
    d = np.ascontiguousarray([int(v) for v in dims], np.int32)
    nV = int(d[0]) * int(d[1]) * int(d[2])
    nN = (int(d[0]) + 1) * (int(d[1]) + 1) * (int(d[2]) + 1)
    scalars = np.zeros(3 * max(1, max_steps))
    counts = np.zeros(2 * max(1, max_steps), np.int32)
    cracked = np.zeros(max(1, max_steps) * nV, np.int32)
    detached = np.zeros(max(1, max_steps) * nV, np.int32)
    vm = np.zeros(max(1, max_steps) * nN, np.float32) if fields else None
    peak, removed, reason = C.c_double(), C.c_int(), C.c_int()
    n = getattr(lib(kind), kind + "_break_test")(d.ctypes.data_as(C.c_void_p), np.ascontiguousarray(density, np.float32).reshape(-1), float(nu),
                                  np.ascontiguousarray(bc, np.uint8).reshape(-1), np.ascontiguousarray(f, np.float64).reshape(-1), float(E),
                                  float(h), float(strength), int(bool(principal)), int(max_steps), float(tol), int(max_iter),
                                  int(coarsest_max_dof), scalars, _ptr(counts), _ptr(cracked), cracked.size, _ptr(detached), detached.size,
                                  _ptr(vm), C.byref(peak), C.byref(removed), C.byref(reason))
    if n < 0:
        raise ValueError(getattr(lib(kind), kind + "_last_error")().decode())
    steps, ic, idt = [], 0, 0
    for s in range(n):
        nc, nd = int(counts[2 * s]), int(counts[2 * s + 1])
        st = {"lambda": scalars[3 * s], "max_disp": scalars[3 * s + 1], "voxels": int(scalars[3 * s + 2]),
              "cracked": cracked[ic:ic + nc].copy(), "detached": detached[idt:idt + nd].copy()}
        if fields:
            st["vm"] = vm[s * nN:(s + 1) * nN].copy()
        steps.append(st)
        ic += nc
        idt += nd
    return {"reason": BREAK_REASONS[reason.value], "peak": peak.value, "removed": removed.value, "steps": steps}


def topo_filter(dims, fill, rmin, x, transpose=False):
    """The density filter (topology.cpp:21-79) over the solid voxels of fill, or its transpose."""
    fill = np.ascontiguousarray(fill, np.float32).reshape(-1)
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    d = np.ascontiguousarray([int(v) for v in dims], np.int32)
    lib("orc").orc_topo_filter(d.ctypes.data_as(C.c_void_p), fill, float(rmin), int(bool(transpose)), x, out)
    return out


def optimize_topology(dims, fill, keep, bc, f, nu=0.3, vol_frac=0.4, penal=3.0, rmin=1.5, move=0.2, tol=0.01, max_iter=40, coarsest_max_dof=1100,
                      kind="orc"):
    """optimizeTopology (topology.cpp:114-183) with the VoxelFEA solve: (density, history rows)."""
    fill = np.ascontiguousarray(fill, np.float32).reshape(-1)
    density = np.zeros_like(fill)
    hist = np.zeros(4 * max(1, int(max_iter)))
    d = np.ascontiguousarray([int(v) for v in dims], np.int32)
    n = getattr(lib(kind), kind + "_optimize_topology")(d.ctypes.data_as(C.c_void_p), fill, np.ascontiguousarray(keep, np.uint8).reshape(-1),
                                                        np.ascontiguousarray(bc, np.uint8).reshape(-1), np.ascontiguousarray(f, np.float64).reshape(-1),
                                                        float(nu), float(vol_frac), float(penal), float(rmin), float(move), float(tol), int(max_iter),
                                                        int(coarsest_max_dof), density, hist)
    if n < 0:
        raise ValueError(getattr(lib(kind), kind + "_last_error")().decode())
    return density, hist[:4 * n].reshape(n, 4)


def grid_for_box(mn, mx, n, pad=0, kind="orc"):
    """gridForBox (voxelize.cpp:14-23): (origin, h, dims)."""
    origin = np.zeros(3)
    h = C.c_double()
    dims = (C.c_int * 3)()
    getattr(lib(kind), kind + "_grid_for_box")(np.ascontiguousarray(mn, np.float64), np.ascontiguousarray(mx, np.float64), int(n), int(pad), origin,
                                                C.byref(h), dims)
    return origin, h.value, tuple(dims)


def voxelize(pos, index, origin, h, dims, sub, kind="orc"):
    """Filled fraction per voxel (voxelize.cpp:25-128), float32 nx * ny * nz."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    index = np.ascontiguousarray(index, np.uint32).reshape(-1)
    frac = np.zeros(int(dims[0]) * int(dims[1]) * int(dims[2]), np.float32)
    d = (C.c_int * 3)(*[int(v) for v in dims])
    r = getattr(lib(kind), kind + "_voxelize")(pos, pos.size // 3, index.ctypes.data_as(C.c_void_p), index.size // 3,
                                              np.ascontiguousarray(origin, np.float64), float(h), d, int(sub), frac)
    if r != 0:
        raise ValueError(getattr(lib(kind), kind + "_last_error")().decode())
    return frac


def principal(s, kind="orc"):
    out = np.zeros(3)
    getattr(lib(kind), kind + "_principal")(*[float(v) for v in s], out)
    return out


def von_mises(s, kind="orc"):
    return getattr(lib(kind), kind + "_von_mises")(np.ascontiguousarray(s, np.float64))


def tridiagonal_max(a, b, kind="orc"):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b if len(b) else [0.0], np.float64)
    return getattr(lib(kind), kind + "_tridiagonal_max")(a, len(a), b)


def cheb_coefficients(lmax, kind="orc"):
    out = np.zeros(3)
    getattr(lib(kind), kind + "_cheb_coefficients")(float(lmax), out)
    return out


def ref_timed_solve(dims, density, nu, bc, f, coarsest_max_dof, tol, max_iter):
    """The reference's own multithreaded build + solve, timed inside the library (CPU baseline)."""
    L = lib("ref")
    u = np.zeros(len(f))
    its, res, tb = C.c_int(), C.c_double(), C.c_double()
    t = L.ref_timed_solve(*[int(d) for d in dims], np.ascontiguousarray(density, np.float32), float(nu),
                          np.ascontiguousarray(bc, np.uint8), int(coarsest_max_dof), np.ascontiguousarray(f, np.float64),
                          float(tol), int(max_iter), u, C.byref(its), C.byref(res), C.byref(tb))
    if t < 0:
        raise ValueError(L.ref_last_error().decode())
    return {"solve_s": t, "build_s": tb.value, "u": u, "iterations": its.value, "residual": res.value}


def ref_timed_phases(dims, density, nu, bc, f, coarsest_max_dof, tol, max_iter):
    """The reference's build, smoother estimates and max_iter PCG iterations, each timed (seconds)."""
    L = lib("ref")
    u = np.zeros(len(f))
    its, res = C.c_int(), C.c_double()
    out = np.zeros(3)
    if L.ref_timed_phases(*[int(d) for d in dims], np.ascontiguousarray(density, np.float32), float(nu), np.ascontiguousarray(bc, np.uint8),
                          int(coarsest_max_dof), np.ascontiguousarray(f, np.float64), float(tol), int(max_iter), u, C.byref(its), C.byref(res), out) < 0:
        raise ValueError(L.ref_last_error().decode())
    return {"build_s": out[0], "smoothers_s": out[1], "solve_s": out[2], "iterations": its.value, "residual": res.value, "u": u}


def ref_threads():
    return int(lib("ref").ref_threads())


# ---- input contract of a CAD part: the reference's own buildPart, typicalWallThickness,
# thinWallDensity, StructuralModel and assemble (oracle/_ref only) ----

def _ref_struct():
    L = lib("ref")
    if not getattr(L, "_struct_set", False):
        i64p = C.POINTER(C.c_int64)
        L.ref_build_part.restype = C.c_int
        L.ref_build_part.argtypes = [_f32p, C.c_int64, np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS"), C.c_int64, i64p, i64p]
        L.ref_part_mesh.restype = None
        L.ref_part_mesh.argtypes = [C.c_int, _f32p, np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS"), _f32p, _f32p]
        L.ref_wall_thickness.restype = C.c_double
        L.ref_wall_thickness.argtypes = [C.c_int, C.c_int]
        L.ref_thin_wall_density.restype = None
        L.ref_thin_wall_density.argtypes = [C.c_int, _f64p, C.c_double, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"), C.c_double, _f32p]
        L.ref_structural_model.restype = C.c_int
        L.ref_structural_model.argtypes = [C.c_int, C.c_int, _f64p, C.POINTER(C.c_double), np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                           C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]
        L.ref_assemble.restype = C.c_int
        L.ref_assemble.argtypes = [C.c_int] + [C.c_void_p] * 4 + [C.c_int, C.c_void_p, C.c_int] + [C.c_void_p] * 6 + [C.c_int, C.c_void_p, C.c_double,
                                   _u8p, _f64p, C.POINTER(C.c_int), _f64p, C.c_char_p, C.c_size_t]
        L._struct_set = True
    return L


def ref_build_part(positions, index):
    """buildPart (weld, orient, refine, triangle data); returns (handle, vertices, tris, tri normals, tri areas)."""
    L = _ref_struct()
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    idx = np.ascontiguousarray(index, np.uint32).reshape(-1)
    nv, nt = C.c_int64(), C.c_int64()
    hnd = L.ref_build_part(pos, pos.size, idx, idx.size, C.byref(nv), C.byref(nt))
    if hnd < 0:
        raise ValueError(L.ref_last_error().decode())
    V, T = np.zeros(3 * nv.value, np.float32), np.zeros(3 * nt.value, np.uint32)
    N, A = np.zeros(3 * nt.value, np.float32), np.zeros(nt.value, np.float32)
    L.ref_part_mesh(hnd, V, T, N, A)
    return hnd, V, T, N, A


def ref_wall_thickness(part, samples=1500):
    return float(_ref_struct().ref_wall_thickness(int(part), int(samples)))


def ref_thin_wall_density(part, origin, h, dims, max_thickness):
    d = np.ascontiguousarray(dims, np.int32)
    out = np.zeros(int(np.prod(d)), np.float32)
    _ref_struct().ref_thin_wall_density(int(part), np.ascontiguousarray(origin, np.float64), float(h), d, float(max_thickness), out)
    return out


def ref_structural_model(part, resolution):
    """StructuralModel(part, resolution): (model handle, origin, h, dims, density, active, surface, wall thickness, (voxels, thin))."""
    L = _ref_struct()
    o, h, d = np.zeros(3), C.c_double(), np.zeros(3, np.int32)
    wall = C.c_double()
    hnd = L.ref_structural_model(int(part), int(resolution), o, C.byref(h), d, None, None, None, C.byref(wall), None)
    if hnd < 0:
        raise ValueError(L.ref_last_error().decode())
    nv, nn = int(np.prod(d)), int(np.prod(d + 1))
    dens, act, surf, cnt = np.zeros(nv, np.float32), np.zeros(nn, np.uint8), np.zeros(nn, np.uint8), np.zeros(2, np.int32)
    hnd = L.ref_structural_model(int(part), int(resolution), o, C.byref(h), d, dens.ctypes.data, act.ctypes.data, surf.ctypes.data, C.byref(wall),
                                 cnt.ctypes.data)
    return hnd, o, h.value, d, dens, act, surf, wall.value, (int(cnt[0]), int(cnt[1]))


def ref_assemble(model, n_nodes, fixtures, loads, gravity, material, to_meters):
    """StructuralModel::assemble with the same fixture / load description parts_sim_amd.assemble takes."""
    L = _ref_struct()
    ntris, tris, has, clip = [], [], [], []

    def add(plist):
        for p in plist:
            t = np.asarray(p["tris"], np.int32)
            ntris.append(t.size)
            tris.append(t)
            c = p.get("clip")
            has.append(1 if c is not None else 0)
            clip.extend(list(c[0]) + [c[1]] if c is not None else [0, 0, 0, 0])
    for pl in fixtures:
        add(pl)
    for l in loads:
        add(l.get("patches", []))
    nt = np.ascontiguousarray(ntris if ntris else [0], np.int64)
    tt = np.ascontiguousarray(np.concatenate(tris) if tris else np.zeros(1, np.int32), np.int32)
    hc = np.ascontiguousarray(has if has else [0], np.int32)
    cl = np.ascontiguousarray(clip if clip else [0.0] * 4, np.float64)
    fnp = np.ascontiguousarray([len(pl) for pl in fixtures] or [0], np.int32)
    types = {"force": 0, "pressure": 1, "wind": 2}
    lt = np.ascontiguousarray([types[l["type"]] for l in loads] or [0], np.int32)
    lm = np.ascontiguousarray([float(l.get("magnitude", 0.0)) for l in loads] or [0.0], np.float64)
    lddir = np.ascontiguousarray([x for l in loads for x in l.get("dir", (0.0, -1.0, 0.0))] or [0.0] * 3, np.float64)
    lnp = np.ascontiguousarray([len(l.get("patches", [])) for l in loads] or [0], np.int32)
    wf = [np.asarray(l["forces"], np.float32).reshape(-1) if l["type"] == "wind" else np.zeros(0, np.float32) for l in loads]
    wn = np.ascontiguousarray([w.size for w in wf] or [0], np.int64)
    wf = np.ascontiguousarray(np.concatenate(wf) if wf and sum(w.size for w in wf) else np.zeros(1, np.float32), np.float32)
    mt = material
    mat = np.ascontiguousarray([mt["E"], mt["nu"], mt["yield"], mt["uts"], mt["density"], float(bool(mt.get("brittle", False)))], np.float64)
    bc, f, tot, fixed = np.zeros(3 * n_nodes, np.uint8), np.zeros(3 * n_nodes), np.zeros(3), C.c_int()
    buf = C.create_string_buffer(8192)
    rc = L.ref_assemble(int(model), nt.ctypes.data, tt.ctypes.data, hc.ctypes.data, cl.ctypes.data, len(fixtures), fnp.ctypes.data, len(loads),
                        lt.ctypes.data, lm.ctypes.data, lddir.ctypes.data, lnp.ctypes.data, wf.ctypes.data, wn.ctypes.data, int(bool(gravity)), mat.ctypes.data,
                        float(to_meters), bc, f, C.byref(fixed), tot, buf, 8192)
    if rc < 0:
        raise ValueError(L.ref_last_error().decode())
    return bc, f, {"fixed_nodes": fixed.value, "total": list(tot)}, [w for w in buf.value.decode().split("\n") if w]
