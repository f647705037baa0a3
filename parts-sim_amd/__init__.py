"""MI355X-native voxel FEA linear-static solve: Python binding of the C ABI in include/psx.h.

The compute path is libpsx.so (HIP kernels for gfx950 + host orchestration). This module only
marshals arrays across the ABI; there is no Python or CPU fallback: without the built library every
call raises. Class and argument names follow the reference's solver boundary
(native/src/gpu/gpu_fea.hpp:17-35, native/src/fea/voxel_fea.hpp:112-189).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PSX_LIB") or os.path.join(HERE, "libpsx.so")

PREC = {"fp64": 0, "ref": 1, "mixed": 2}
OK, ERR_ARGUMENT, ERR_NO_DEVICE, ERR_NO_COARSE, ERR_DEVICE, ERR_INTERNAL = range(6)


class PsxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("psx error %d: %s" % (code, msg))
        self.code = code


class ModelDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("density", C.c_void_p), ("nu", C.c_double),
                ("bc", C.c_void_p), ("diag_add", C.c_void_p), ("coarsest_max_dof", C.c_int64), ("max_levels", C.c_int),
                ("build_on_device", C.c_int)]


PROGRESS = C.CFUNCTYPE(C.c_int, C.c_int, C.c_double, C.c_void_p)


class SolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("precision", C.c_int), ("check_every", C.c_int),
                ("fold_every", C.c_int), ("on_progress", PROGRESS), ("user", C.c_void_p), ("cheb_degree", C.c_int),
                ("mfma_smoother", C.c_int), ("eager", C.c_int), ("pre_degree", C.c_int)]


TOPO_ITER = C.CFUNCTYPE(C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_float), C.c_void_p)


class TopologyOpts(C.Structure):
    _fields_ = [("vol_frac", C.c_double), ("penal", C.c_double), ("rmin", C.c_double), ("move", C.c_double), ("tol", C.c_double),
                ("max_iter", C.c_int), ("precision", C.c_int), ("solve_tol", C.c_double), ("coarsest_max_dof", C.c_int64),
                ("on_iter", TOPO_ITER), ("user", C.c_void_p), ("solve_max_iter", C.c_int)]


class BreakStep(C.Structure):
    _fields_ = [("step", C.c_int), ("lambda_", C.c_double), ("max_disp", C.c_double), ("voxels", C.c_int),
                ("cracked", C.POINTER(C.c_int32)), ("n_cracked", C.c_int), ("detached", C.POINTER(C.c_int32)), ("n_detached", C.c_int),
                ("u", C.POINTER(C.c_float)), ("node_vm", C.POINTER(C.c_float)), ("active_node", C.POINTER(C.c_uint8))]


BREAK_FN = C.CFUNCTYPE(C.c_int, C.POINTER(BreakStep), C.c_void_p)


class BreakOpts(C.Structure):
    _fields_ = [("strength", C.c_double), ("principal", C.c_int), ("max_steps", C.c_int), ("E", C.c_double), ("h", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("precision", C.c_int), ("tie_order", C.c_int), ("fields", C.c_int),
                ("on_step", BREAK_FN), ("user", C.c_void_p)]


class BreakResult(C.Structure):
    _fields_ = [("reason", C.c_int), ("steps", C.c_int), ("removed", C.c_int), ("peak", C.c_double)]


BREAK_REASONS = ("step limit", "separated", "unsupported load", "solver did not converge", "no stress", "stopped")


class ServerOpts(C.Structure):
    _fields_ = [("max_inflight", C.c_int), ("hbm_budget", C.c_int64), ("pack_max", C.c_int), ("pack_max_dof", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("part", ModelDesc), ("nrhs", C.c_int), ("f", C.c_void_p), ("x0", C.c_void_p), ("x0_per_rhs", C.c_int),
                ("part_id", C.c_int64), ("opts", SolveOpts)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("max_iterations", C.c_int), ("converged", C.c_int), ("residual", C.c_double),
                ("queue_ms", C.c_double), ("build_ms", C.c_double), ("solve_ms", C.c_double), ("total_ms", C.c_double),
                ("device_bytes", C.c_int64), ("reused_part", C.c_int), ("packed", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("converged", C.c_int), ("cancelled", C.c_int),
                ("broken", C.c_int), ("folds", C.c_int), ("device_ms", C.c_double)]


class LevelInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n_nodes", C.c_int64), ("n_dof", C.c_int64),
                ("free_dof", C.c_int64), ("n_elems", C.c_int64), ("n_own", C.c_int64), ("lmax", C.c_double)]


class PsxMesh(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("n_vertices", C.c_int64), ("tris", C.c_void_p), ("n_tris", C.c_int64)]


class PsxPatch(C.Structure):
    _fields_ = [("tris", C.c_void_p), ("n_tris", C.c_int64), ("has_clip", C.c_int), ("clip_center", C.c_double * 3), ("clip_radius", C.c_double)]


class PsxFixture(C.Structure):
    _fields_ = [("name", C.c_char_p), ("patches", C.POINTER(PsxPatch)), ("n_patches", C.c_int)]


class PsxLoad(C.Structure):
    _fields_ = [("type", C.c_int), ("name", C.c_char_p), ("magnitude", C.c_double), ("dir", C.c_double * 3), ("patches", C.POINTER(PsxPatch)),
                ("n_patches", C.c_int), ("forces", C.c_void_p), ("n_forces", C.c_int64)]


class PsxMaterial(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("yield_", C.c_double), ("uts", C.c_double), ("density", C.c_double), ("brittle", C.c_int)]


class StructuralInfo(C.Structure):
    _fields_ = [("wall_thickness", C.c_double), ("sub", C.c_int), ("voxel_count", C.c_int), ("thin_voxels", C.c_int), ("used_shell", C.c_int)]


class AssemblyInfo(C.Structure):
    _fields_ = [("fixed_nodes", C.c_int), ("total", C.c_double * 3), ("n_warnings", C.c_int)]


LOAD_TYPES = {"force": 0, "pressure": 1, "wind": 2}

# every entry point declared in include/psx.h: name -> (restype, argtypes)
_vp, _i, _d, _i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
SIGNATURES = {
    "psx_last_error": (C.c_char_p, []),
    "psx_available": (_i, []),
    "psx_device_name": (_i, [C.c_char_p, C.c_size_t]),
    "psx_create": (_i, [C.POINTER(ModelDesc), C.POINTER(_vp)]),
    "psx_destroy": (None, [_vp]),
    "psx_num_levels": (_i, [_vp]),
    "psx_get_level_info": (_i, [_vp, _i, C.POINTER(LevelInfo)]),
    "psx_estimate_smoothers": (_i, [_vp]),
    "psx_set_lmax": (_i, [_vp, _i, _d]),
    "psx_apply": (_i, [_vp, _i, _vp, _vp, _i]),
    "psx_apply32": (_i, [_vp, _i, _vp, _vp, _i]),
    "psx_apply_mfma": (_i, [_vp, _i, _i, _vp, _vp]),
    "psx_apply_child": (_i, [_vp, _i, _i, _vp, _vp, _i]),
    "psx_precondition": (_i, [_vp, _vp, _vp]),
    "psx_precondition_device": (_i, [_vp, _vp, _vp]),
    "psx_precondition_multi": (_i, [_vp, _i, _vp, _vp]),
    "psx_precondition_multi_device": (_i, [_vp, _i, _vp, _vp]),
    "psx_solve": (_i, [_vp, _vp, _vp, C.POINTER(SolveOpts), _vp, C.POINTER(SolveInfo)]),
    "psx_solve_device": (_i, [_vp, _vp, _vp, C.POINTER(SolveOpts), _vp, C.POINTER(SolveInfo)]),
    "psx_solve_multi": (_i, [_vp, _i, _vp, _vp, C.POINTER(SolveOpts), _vp, C.POINTER(SolveInfo)]),
    "psx_solve_multi_device": (_i, [_vp, _i, _vp, _vp, C.POINTER(SolveOpts), _vp, C.POINTER(SolveInfo)]),
    "psx_solve_packed": (_i, [_i, _vp, _vp, _vp, _vp, C.POINTER(SolveOpts), _vp, C.POINTER(SolveInfo)]),
    "psx_reactions": (_i, [_vp, _vp, _vp, _vp]),
    "psx_stresses": (_i, [_vp, _vp, _d, _d, _i, _vp, _vp, _vp]),
    "psx_stresses_device": (_i, [_vp, _vp, _d, _d, _i, _vp, _vp, _vp]),
    "psx_stress_tensors": (_i, [_vp, _vp, _d, _d, _vp]),
    "psx_profile": (_i, [_vp, _i, C.c_char_p, C.c_size_t]),
    "psx_device_bytes": (_i64, [_vp]),
    "psx_server_create": (_i, [C.POINTER(ServerOpts), C.POINTER(_vp)]),
    "psx_server_submit": (_i, [_vp, C.POINTER(Request), C.POINTER(_i64)]),
    "psx_server_wait": (_i, [_vp, _i64, _vp, C.POINTER(Result)]),
    "psx_server_pending": (_i, [_vp]),
    "psx_server_destroy": (None, [_vp]),
    "psx_prune_floating": (_i, [_i, _i, _i, _vp, _vp, _vp]),
    "psx_grid_for_box": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    "psx_voxelize": (_i, [_vp, C.c_int64, _vp, C.c_int64, _vp, C.c_double, _vp, _i, _vp]),
    "psx_optimize_topology": (_i, [_vp, _vp, _vp, _vp, _vp, _d, C.POINTER(TopologyOpts), _vp, _vp, C.POINTER(_i), C.POINTER(_d)]),
    "psx_element_energies": (_i, [_vp, _vp, _vp]),
    "psx_topology_filter": (_i, [_vp, _vp, _d, _i, _vp, _vp]),
    "psx_break_test": (_i, [C.POINTER(ModelDesc), _vp, C.POINTER(BreakOpts), C.POINTER(BreakResult)]),
    "psx_prune_floating_device": (_i, [_i, _i, _i, _vp, _vp, _vp, C.POINTER(_i)]),
    "psx_prune_floating_device_noalloc": (_i, [_i, _i, _i, _vp, _vp, _vp, C.POINTER(_i)]),
    "psx_hex_element": (_i, [_d, _vp, _vp]),
    "psx_hierarchy_create": (_i, [C.POINTER(ModelDesc), C.POINTER(_vp)]),
    "psx_hierarchy_destroy": (None, [_vp]),
    "psx_hierarchy_num_levels": (_i, [_vp]),
    "psx_hierarchy_level_info": (_i, [_vp, _i, C.POINTER(LevelInfo)]),
    "psx_hierarchy_level": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "psx_model_level": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "psx_model_level_child": (_i, [_vp, _i, _vp, _vp, _vp]),
    "psx_model_coarse": (_i, [_vp, C.POINTER(C.c_int64), _vp, _vp]),
    "psx_hierarchy_coarse": (_i, [_vp, C.POINTER(_i64), _vp, _vp]),
    "psx_wall_thickness": (_i, [C.POINTER(PsxMesh), _i, C.POINTER(_d)]),
    "psx_thin_wall_density": (_i, [C.POINTER(PsxMesh), _vp, _d, _vp, _d, _vp]),
    "psx_structural_model": (_i, [C.POINTER(PsxMesh), _vp, _d, _vp, _vp, _vp, _vp, C.POINTER(StructuralInfo)]),
    "psx_assemble": (_i, [C.POINTER(PsxMesh), _vp, _d, _vp, _vp, C.POINTER(PsxFixture), _i, C.POINTER(PsxLoad), _i, _i, C.POINTER(PsxMaterial), _d,
                          _vp, _vp, C.POINTER(AssemblyInfo), C.c_char_p, C.c_size_t]),
}

_lib = None


def lib():
    """The loaded libpsx.so (raises when it is not built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(LIB_PATH + " is not built: run make -C parts-sim_amd (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(L, name):  # an older build: the entry point raises when called
                continue
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(code):
    if code != OK:
        raise PsxError(code, lib().psx_last_error().decode())


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def available():
    return bool(lib().psx_available())


def device_name():
    buf = C.create_string_buffer(256)
    _check(lib().psx_device_name(buf, len(buf)))
    return buf.value.decode()


def _desc(dims, density, nu, bc, coarsest_max_dof, diag_add, max_levels, keep, device_build=False):
    density = np.ascontiguousarray(density, np.float32).reshape(-1)
    bc = np.ascontiguousarray(bc, np.uint8).reshape(-1)
    nx, ny, nz = (int(v) for v in dims)
    if density.size != nx * ny * nz or bc.size != 3 * (nx + 1) * (ny + 1) * (nz + 1):
        raise ValueError("Structural grid, density and fixture sizes do not match.")
    da = None if diag_add is None else np.ascontiguousarray(diag_add, np.float64)
    keep.extend([density, bc, da])
    return ModelDesc(nx, ny, nz, density.ctypes.data, float(nu), bc.ctypes.data, None if da is None else da.ctypes.data,
                     int(coarsest_max_dof), int(max_levels), int(bool(device_build)))


def _level_info(fn, h, l):
    li = LevelInfo()
    _check(fn(h, l, C.byref(li)))
    return {"dims": (li.nx, li.ny, li.nz), "nNodes": li.n_nodes, "nDof": li.n_dof, "freeDof": li.free_dof,
            "nElems": li.n_elems, "nOwn": li.n_own, "lmax": li.lmax}


def prune_floating(dims, density, held):
    """Density with the voxels not face-connected to a held node removed; (density, removed)."""
    nx, ny, nz = (int(v) for v in dims)
    d = np.ascontiguousarray(density, np.float32).reshape(-1)
    h = np.ascontiguousarray(held, np.uint8).reshape(-1)
    if d.size != nx * ny * nz or h.size != (nx + 1) * (ny + 1) * (nz + 1):
        raise ValueError("Structural grid, density and fixture sizes do not match.")
    out = np.zeros_like(d)
    r = lib().psx_prune_floating(nx, ny, nz, _p(d), _p(h), _p(out))
    if r < 0:
        raise PsxError(-r, lib().psx_last_error().decode())
    return out, r


def grid_for_box(mn, mx, n, pad=0):
    """Grid enclosing the box [mn, mx] with n voxels along its longest side (gridForBox): (origin, h, dims)."""
    mn = np.ascontiguousarray(mn, np.float64)
    mx = np.ascontiguousarray(mx, np.float64)
    origin = np.zeros(3)
    h = C.c_double()
    dims = np.zeros(3, np.int32)
    _check(lib().psx_grid_for_box(_p(mn), _p(mx), int(n), int(pad), _p(origin), C.byref(h), _p(dims)))
    return origin, h.value, tuple(int(v) for v in dims)


def voxelize(pos, index, origin, h, dims, sub=2):
    """Filled fraction per voxel of a triangle mesh, computed on the device (float32, i + nx (j + ny k))."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    index = np.ascontiguousarray(index, np.uint32).reshape(-1)
    if pos.size % 3 or index.size % 3:
        raise ValueError("Positions and triangle indices come in threes.")
    d = np.ascontiguousarray([int(v) for v in dims], np.int32)
    o = np.ascontiguousarray(origin, np.float64)
    frac = np.zeros(int(d[0]) * int(d[1]) * int(d[2]), np.float32)
    _check(lib().psx_voxelize(_p(pos), pos.size // 3, _p(index), index.size // 3, _p(o), float(h), _p(d), int(sub), _p(frac)))
    return frac


def _mesh(vertices, tris, keep):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
    t = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    if v.size % 3 or t.size % 3:
        raise ValueError("Positions and triangle indices come in threes.")
    keep.extend([v, t])
    return PsxMesh(v.ctypes.data, v.size // 3, t.ctypes.data, t.size // 3)


def wall_thickness(vertices, tris, samples=1500):
    """typicalWallThickness of a surface mesh: the 20th percentile of the sampled wall thickness, or inf."""
    keep, out = [], C.c_double()
    m = _mesh(vertices, tris, keep)
    _check(lib().psx_wall_thickness(C.byref(m), int(samples), C.byref(out)))
    return out.value


def thin_wall_density(vertices, tris, origin, h, dims, max_thickness):
    """thinWallDensity: the density thin walls give each voxel of the grid (float32, 0 where none)."""
    keep = []
    m = _mesh(vertices, tris, keep)
    d = np.ascontiguousarray([int(x) for x in dims], np.int32)
    o = np.ascontiguousarray(origin, np.float64)
    out = np.zeros(int(np.prod(d)), np.float32)
    _check(lib().psx_thin_wall_density(C.byref(m), _p(o), float(h), _p(d), float(max_thickness), _p(out)))
    return out


def structural_model(vertices, tris, origin, h, dims):
    """StructuralModel's voxel densities, active and surface nodes and counts on the given grid (the
    fill fractions from the device voxelizer)."""
    keep = []
    m = _mesh(vertices, tris, keep)
    d = np.ascontiguousarray([int(x) for x in dims], np.int32)
    o = np.ascontiguousarray(origin, np.float64)
    dens = np.zeros(int(np.prod(d)), np.float32)
    nn = int(np.prod(d + 1))
    act, surf = np.zeros(nn, np.uint8), np.zeros(nn, np.uint8)
    info = StructuralInfo()
    _check(lib().psx_structural_model(C.byref(m), _p(o), float(h), _p(d), _p(dens), _p(act), _p(surf), C.byref(info)))
    return dens, act, surf, {k: getattr(info, k) for k, _ in StructuralInfo._fields_}


def _patches(plist, keep):
    arr = (PsxPatch * max(1, len(plist)))()
    for q, p in enumerate(plist):
        t = np.ascontiguousarray(p["tris"], np.int32)
        keep.append(t)
        arr[q].tris, arr[q].n_tris = t.ctypes.data, t.size
        if p.get("clip") is not None:
            c, r = p["clip"]
            arr[q].has_clip = 1
            arr[q].clip_center = (C.c_double * 3)(*[float(x) for x in c])
            arr[q].clip_radius = float(r)
    keep.append(arr)
    return arr


def assemble(vertices, tris, origin, h, dims, density, fixtures, loads, gravity=False, material=None, to_meters=1e-3):
    """StructuralModel::assemble: held DOFs and nodal forces of fixtures (lists of patches) and loads
    (dicts: type force / pressure / wind, magnitude, dir, patches, forces) on the grid's nodes.
    A patch is {"tris": triangle indices, "clip": None or ((x, y, z), radius)}. material: dict with
    E, nu, yield, uts, density, brittle. Returns (bc, f, info, warnings)."""
    keep = []
    m = _mesh(vertices, tris, keep)
    d = np.ascontiguousarray([int(x) for x in dims], np.int32)
    o = np.ascontiguousarray(origin, np.float64)
    dens = np.ascontiguousarray(density, np.float32).reshape(-1)
    fx = (PsxFixture * max(1, len(fixtures)))()
    for q, pl in enumerate(fixtures):
        fx[q].name = b"F%d" % q
        fx[q].patches = _patches(pl, keep)
        fx[q].n_patches = len(pl)
    ld = (PsxLoad * max(1, len(loads)))()
    for q, l in enumerate(loads):
        ld[q].type = LOAD_TYPES[l["type"]]
        ld[q].name = b"L%d" % q
        ld[q].magnitude = float(l.get("magnitude", 0.0))
        ld[q].dir = (C.c_double * 3)(*[float(x) for x in l.get("dir", (0.0, -1.0, 0.0))])
        ld[q].patches = _patches(l.get("patches", []), keep)
        ld[q].n_patches = len(l.get("patches", []))
        if l.get("forces") is not None:
            fw = np.ascontiguousarray(l["forces"], np.float32).reshape(-1)
            keep.append(fw)
            ld[q].forces, ld[q].n_forces = fw.ctypes.data, fw.size
    mt = material or {"E": 200.0, "nu": 0.3, "yield": 250.0, "uts": 400.0, "density": 7850.0, "brittle": False}
    mat = PsxMaterial(mt["E"], mt["nu"], mt["yield"], mt["uts"], mt["density"], int(bool(mt.get("brittle", False))))
    nn = int(np.prod(d + 1))
    bc, f = np.zeros(3 * nn, np.uint8), np.zeros(3 * nn)
    info = AssemblyInfo()
    buf = C.create_string_buffer(8192)
    _check(lib().psx_assemble(C.byref(m), _p(o), float(h), _p(d), _p(dens), fx, len(fixtures), ld, len(loads), int(bool(gravity)), C.byref(mat),
                              float(to_meters), _p(bc), _p(f), C.byref(info), buf, 8192))
    out = {"fixed_nodes": info.fixed_nodes, "total": list(info.total), "n_warnings": info.n_warnings}
    return bc, f, out, [w for w in buf.value.decode().split("\n") if w]


def optimize_topology(dims, fill, keep, bc, f, nu, vol_frac=0.4, penal=3.0, rmin=1.5, move=0.2, tol=0.01, max_iter=40, precision="mixed",
                      solve_tol=1e-8, coarsest_max_dof=1100, on_iter=None, solve_max_iter=1000):
    """SIMP topology optimization (optimizeTopology): the static solves, energies, filter and OC bisection
    on the device. Returns (density per voxel, history rows [it, compliance, volume, change], kept fraction).
    on_iter(it, compliance, volume, change, density) returning True stops the loop (density: the iteration's
    filtered densities per voxel, a copy)."""
    d = np.ascontiguousarray([int(v) for v in dims], np.int32)
    nV = int(d[0]) * int(d[1]) * int(d[2])
    fill = np.ascontiguousarray(fill, np.float32).reshape(-1)
    keep = np.ascontiguousarray(keep, np.uint8).reshape(-1)
    bc = np.ascontiguousarray(bc, np.uint8).reshape(-1)
    f = np.ascontiguousarray(f, np.float64).reshape(-1)
    nDof = 3 * (int(d[0]) + 1) * (int(d[1]) + 1) * (int(d[2]) + 1)
    if fill.size != nV or keep.size != nV or bc.size != nDof or f.size != nDof:
        raise ValueError("fill and keep need one value per voxel, bc and f one per DOF")
    o = TopologyOpts(vol_frac=vol_frac, penal=penal, rmin=rmin, move=move, tol=tol, max_iter=max_iter, precision=PREC[precision],
                     solve_tol=solve_tol, coarsest_max_dof=coarsest_max_dof, solve_max_iter=solve_max_iter)
    cb = None
    if on_iter is not None:
        cb = TOPO_ITER(lambda it, c, v, ch, dens, _u: 1 if on_iter(it, c, v, ch, np.ctypeslib.as_array(dens, (nV,)).copy()) else 0)
        o.on_iter = cb
    density = np.zeros(nV, np.float32)
    history = np.zeros(4 * max(1, int(max_iter)))
    iters, kept = C.c_int(), C.c_double()
    _check(lib().psx_optimize_topology(_p(d), _p(fill), _p(keep), _p(bc), _p(f), float(nu), C.byref(o), _p(density), _p(history),
                                       C.byref(iters), C.byref(kept)))
    return density, history[:4 * iters.value].reshape(-1, 4), kept.value


def topology_filter(dims, fill, rmin, x, transpose=False):
    """The density filter (or its transpose) over the voxels with fill > 0, on the device."""
    d = np.ascontiguousarray([int(v) for v in dims], np.int32)
    fill = np.ascontiguousarray(fill, np.float32).reshape(-1)
    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    if x.size != int((fill > 0).sum()):
        raise ValueError("x needs one value per voxel with fill > 0")
    out = np.zeros_like(x)
    _check(lib().psx_topology_filter(_p(d), _p(fill), float(rmin), int(bool(transpose)), _p(x), _p(out)))
    return out


def break_test(dims, density, nu, bc, f, E, h, strength, principal=False, max_steps=5, tol=1e-6, max_iter=3000, precision="mixed",
               tie_order=0, fields=False, coarsest_max_dof=1100, on_step=None):
    """The break test (breakTest) with the solves, stresses, measures and prunes on the device: dict(reason,
    peak, removed, steps), each step dict(lambda, max_disp, voxels, cracked, detached[, u, vm, active]).
    on_step(step_dict) returning True stops the sequence."""
    dims = tuple(int(d) for d in dims)
    density = np.ascontiguousarray(density, np.float32).reshape(-1)
    bc = np.ascontiguousarray(bc, np.uint8).reshape(-1)
    f = np.ascontiguousarray(f, np.float64).reshape(-1)
    nN = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1)
    if density.size != dims[0] * dims[1] * dims[2] or bc.size != 3 * nN or f.size != 3 * nN:
        raise ValueError("density needs one value per voxel, bc and f one per DOF")
    desc = ModelDesc(nx=dims[0], ny=dims[1], nz=dims[2], density=density.ctypes.data, nu=float(nu), bc=bc.ctypes.data, diag_add=None,
                     coarsest_max_dof=int(coarsest_max_dof), max_levels=0)
    steps = []

    def collect(ptr, _user):
        b = ptr.contents
        st = {"step": b.step, "lambda": b.lambda_, "max_disp": b.max_disp, "voxels": b.voxels,
              "cracked": np.ctypeslib.as_array(b.cracked, (b.n_cracked,)).copy() if b.n_cracked else np.zeros(0, np.int32),
              "detached": np.ctypeslib.as_array(b.detached, (b.n_detached,)).copy() if b.n_detached else np.zeros(0, np.int32)}
        if fields:
            st["u"] = np.ctypeslib.as_array(b.u, (3 * nN,)).copy()
            st["vm"] = np.ctypeslib.as_array(b.node_vm, (nN,)).copy()
            st["active"] = np.ctypeslib.as_array(b.active_node, (nN,)).copy()
        steps.append(st)
        return 1 if (on_step is not None and on_step(st)) else 0

    cb = BREAK_FN(collect)
    o = BreakOpts(strength=float(strength), principal=int(bool(principal)), max_steps=int(max_steps), E=float(E), h=float(h), tol=float(tol),
                  max_iter=int(max_iter), precision=PREC[precision], tie_order=int(tie_order), fields=int(bool(fields)), on_step=cb, user=None)
    res = BreakResult()
    _check(lib().psx_break_test(C.byref(desc), _p(f), C.byref(o), C.byref(res)))
    return {"reason": BREAK_REASONS[res.reason], "peak": res.peak, "removed": res.removed, "steps": steps}


def hex_element(nu):
    K = np.zeros(576)
    cs = np.zeros(8 * 144)
    _check(lib().psx_hex_element(float(nu), _p(K), _p(cs)))
    return K.reshape(24, 24), cs.reshape(8, 6, 24)


class Hierarchy:
    """The multigrid hierarchy built on the host, exactly as a model uploads it (no device)."""

    def __init__(self, dims, density, nu, bc, coarsest_max_dof=1100, diag_add=None, max_levels=0):
        keep = []
        d = _desc(dims, density, nu, bc, coarsest_max_dof, diag_add, max_levels, keep)
        h = C.c_void_p()
        _check(lib().psx_hierarchy_create(C.byref(d), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().psx_hierarchy_destroy(self.h)
            self.h = None

    @property
    def num_levels(self):
        return lib().psx_hierarchy_num_levels(self.h)

    def level(self, l):
        return _level_info(lib().psx_hierarchy_level_info, self.h, l)

    def level_arrays(self, l):
        info = self.level(l)
        nvox = int(np.prod(info["dims"]))
        fixed = np.zeros(info["nDof"], np.uint8)
        inv = np.zeros(info["nDof"])
        ns = np.zeros(info["nNodes"], np.float32)
        words = np.zeros(nvox, np.int32)
        kb = np.zeros(576)
        own = np.zeros(max(1, info["nOwn"]) * 576, np.float32)
        _check(lib().psx_hierarchy_level(self.h, l, _p(fixed), _p(inv), _p(ns), _p(words), _p(kb), _p(own)))
        return {"fixed": fixed, "invDiag": inv, "nodeScale": ns, "words": words, "Kb": kb.reshape(24, 24),
                "own": own[:info["nOwn"] * 576].reshape(-1, 24, 24)}

    def coarse(self):
        m = C.c_int64()
        _check(lib().psx_hierarchy_coarse(self.h, C.byref(m), None, None))
        idx = np.zeros(max(1, m.value), np.int32)
        ainv = np.zeros(max(1, m.value * m.value))
        _check(lib().psx_hierarchy_coarse(self.h, C.byref(m), _p(idx), _p(ainv)))
        return idx[:m.value], ainv[:m.value * m.value].reshape(m.value, m.value)


class Model:
    """One part on the GPU: its hierarchy in HBM and the solver around it
    (VoxelFEA + GpuFeaSolver of the reference)."""

    def __init__(self, dims, density, nu, bc, coarsest_max_dof=1100, diag_add=None, max_levels=0, device_build=False):
        """device_build: build the hierarchy with the device kernels (the host build's level data bit
        for bit) instead of on the host CPU."""
        keep = []
        d = _desc(dims, density, nu, bc, coarsest_max_dof, diag_add, max_levels, keep, device_build)
        h = C.c_void_p()
        _check(lib().psx_create(C.byref(d), C.byref(h)))
        self.h = h
        self.dims = tuple(int(v) for v in dims)
        self.n_dof = self.level(0)["nDof"]

    def close(self):
        if getattr(self, "h", None):
            lib().psx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def num_levels(self):
        return lib().psx_num_levels(self.h)

    def level(self, l):
        return _level_info(lib().psx_get_level_info, self.h, l)

    def level_arrays(self, l):
        """The level data the kernels read, downloaded from the device (the host Hierarchy's fields,
        plus the node info words, the child-scale form and the diagonal shift where the level has them)."""
        info = self.level(l)
        nvox = int(np.prod(info["dims"]))
        fixed = np.zeros(info["nDof"], np.uint8)
        inv = np.zeros(info["nDof"])
        ns = np.zeros(info["nNodes"], np.float32)
        words = np.zeros(nvox, np.int32)
        kb = np.zeros(576)
        own = np.zeros(max(1, info["nOwn"]) * 576, np.float32)
        ninfo = np.zeros(info["nNodes"], np.uint16)
        _check(lib().psx_model_level(self.h, l, _p(fixed), _p(inv), _p(ns), _p(words), _p(kb), _p(own), _p(ninfo)))
        out = {"fixed": fixed, "invDiag": inv, "nodeScale": ns, "words": words, "Kb": kb.reshape(24, 24),
               "own": own[:info["nOwn"] * 576].reshape(-1, 24, 24), "nodeInfo": ninfo}
        cs = np.zeros(max(1, info["nOwn"]) * 8, np.float32)
        two = np.zeros(max(1, info["nOwn"]) * 4, np.int32)
        if lib().psx_model_level_child(self.h, l, _p(cs), _p(two), None) == 0:
            out["childS"], out["childTwo"] = cs[:info["nOwn"] * 8], two[:info["nOwn"] * 4]
        da = np.zeros(info["nDof"])
        if lib().psx_model_level_child(self.h, l, None, None, _p(da)) == 0:
            out["diagAdd"] = da
        return out

    def coarse(self):
        """The coarsest level's free-DOF map and dense inverse as the device holds them."""
        m = C.c_int64()
        _check(lib().psx_model_coarse(self.h, C.byref(m), None, None))
        idx = np.zeros(max(1, m.value), np.int32)
        ainv = np.zeros(max(1, m.value * m.value))
        _check(lib().psx_model_coarse(self.h, C.byref(m), _p(idx), _p(ainv)))
        return idx[:m.value], ainv[:m.value * m.value].reshape(m.value, m.value)

    def estimate_smoothers(self):
        _check(lib().psx_estimate_smoothers(self.h))
        return [self.level(l)["lmax"] for l in range(self.num_levels)]

    def set_lmax(self, level, lmax):
        _check(lib().psx_set_lmax(self.h, int(level), float(lmax)))

    def apply(self, x, level=0, raw=False):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        _check(lib().psx_apply(self.h, int(level), _p(x), _p(y), int(raw)))
        return y

    def apply32(self, x, level=0, raw=False):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros_like(x)
        _check(lib().psx_apply32(self.h, int(level), _p(x), _p(y), int(raw)))
        return y

    def apply_child(self, x, level=0, raw=False):
        """y = K x as the FP64 and mixed solves apply it (own matrices in child-scale form); float32 x
        gives the FP32 operator."""
        fp32 = np.asarray(x).dtype == np.float32
        x = np.ascontiguousarray(x, np.float32 if fp32 else np.float64)
        y = np.zeros_like(x)
        _check(lib().psx_apply_child(self.h, int(level), int(fp32), _p(x), _p(y), int(raw)))
        return y

    def apply_mfma(self, x, level=0, pieces=2):
        """y = K x (FP32) as the element-by-element matrix-core product with `pieces` BF16 pieces of the
        operands (2: the mixed V-cycle's smoothing; 3: the outer operator's option)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros_like(x)
        _check(lib().psx_apply_mfma(self.h, int(level), int(pieces), _p(x), _p(y)))
        return y

    def precondition(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros_like(r)
        _check(lib().psx_precondition(self.h, _p(r), _p(z)))
        return z

    def precondition_multi(self, R):
        """One V-cycle per row of R (nvec x n_dof) in one set of launches."""
        R = np.ascontiguousarray(R, np.float64)
        if R.ndim != 2 or R.shape[1] != self.n_dof:
            raise ValueError("R must be nvec x n_dof")
        Z = np.zeros_like(R)
        _check(lib().psx_precondition_multi(self.h, R.shape[0], _p(R), _p(Z)))
        return Z

    @staticmethod
    def _opts(tol, max_iter, precision, check_every, fold_every, on_progress, cheb_degree=0, mfma_smoother=False, eager=False, pre_degree=0):
        cb = PROGRESS(0) if on_progress is None else PROGRESS(lambda it, rel, user: int(bool(on_progress(it, rel))))
        return SolveOpts(float(tol), int(max_iter), PREC[precision], int(check_every), int(fold_every), cb, None, int(cheb_degree),
                         int(mfma_smoother), int(bool(eager)), int(pre_degree)), cb

    @staticmethod
    def _info(info):
        return {"iterations": info.iterations, "residual": info.residual, "converged": bool(info.converged),
                "cancelled": bool(info.cancelled), "broken": bool(info.broken), "folds": info.folds, "device_ms": info.device_ms}

    def solve(self, f, tol=1e-6, max_iter=500, x0=None, precision="fp64", check_every=0, fold_every=0, on_progress=None, cheb_degree=0,
              mfma_smoother=False, eager=False, pre_degree=0):
        f = np.ascontiguousarray(f, np.float64)
        if f.size != self.n_dof or (x0 is not None and np.size(x0) != self.n_dof):
            raise ValueError("Force or displacement vector has the wrong size.")
        x0a = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        u = np.zeros_like(f)
        o, cb = self._opts(tol, max_iter, precision, check_every, fold_every, on_progress, cheb_degree, mfma_smoother, eager, pre_degree)
        info = SolveInfo()
        _check(lib().psx_solve(self.h, _p(f), _p(x0a), C.byref(o), _p(u), C.byref(info)))
        out = self._info(info)
        out["u"] = u
        return out

    def solve_device(self, d_f, d_u, d_x0=None, tol=1e-6, max_iter=500, precision="fp64", check_every=0, fold_every=0, cheb_degree=0,
                     mfma_smoother=False, eager=False, pre_degree=0):
        """Solve with HBM-resident vectors (DeviceArray or raw device addresses)."""
        o, cb = self._opts(tol, max_iter, precision, check_every, fold_every, None, cheb_degree, mfma_smoother, eager, pre_degree)
        info = SolveInfo()
        addr = lambda a: None if a is None else C.c_void_p(int(a))
        _check(lib().psx_solve_device(self.h, addr(d_f), addr(d_x0), C.byref(o), addr(d_u), C.byref(info)))
        return self._info(info)

    def solve_multi(self, F, tol=1e-6, max_iter=500, X0=None, precision="mixed", check_every=0, fold_every=0, on_progress=None, cheb_degree=0,
                    eager=False, mfma_smoother=0):
        """Solve K U[k] = F[k] for nrhs load cases (F: nrhs x n_dof) in one set of launches per
        V-cycle level; returns U and one info dict per load case."""
        F = np.ascontiguousarray(F, np.float64)
        if F.ndim != 2 or F.shape[1] != self.n_dof or F.shape[0] < 1:
            raise ValueError("Force matrix must be nrhs x n_dof.")
        k = F.shape[0]
        X0a = None if X0 is None else np.ascontiguousarray(X0, np.float64)
        if X0a is not None and X0a.size != F.size:
            raise ValueError("Start displacements must be nrhs x n_dof.")
        U = np.zeros_like(F)
        o, cb = self._opts(tol, max_iter, precision, check_every, fold_every, on_progress, cheb_degree, mfma_smoother, eager)
        infos = (SolveInfo * k)()
        _check(lib().psx_solve_multi(self.h, k, _p(F), _p(X0a), C.byref(o), _p(U), infos))
        return U, [self._info(infos[i]) for i in range(k)]

    def solve_multi_device(self, nrhs, d_f, d_u, d_x0=None, tol=1e-6, max_iter=500, precision="mixed", check_every=0, fold_every=0, cheb_degree=0,
                           eager=False, mfma_smoother=0):
        """solve_multi with HBM-resident nrhs x n_dof vectors (DeviceArray or raw device addresses)."""
        o, cb = self._opts(tol, max_iter, precision, check_every, fold_every, None, cheb_degree, mfma_smoother, eager)
        infos = (SolveInfo * int(nrhs))()
        addr = lambda a: None if a is None else C.c_void_p(int(a))
        _check(lib().psx_solve_multi_device(self.h, int(nrhs), addr(d_f), addr(d_x0), C.byref(o), addr(d_u), infos))
        return [self._info(infos[i]) for i in range(int(nrhs))]

    def reactions(self, u, f=None):
        R = np.zeros(3)
        fa = None if f is None else np.ascontiguousarray(f, np.float64)
        _check(lib().psx_reactions(self.h, _p(np.ascontiguousarray(u, np.float64)), _p(fa), _p(R)))
        return R

    def element_energies(self, u):
        """u_e^T K0 u_e of the finest level's elements in voxel order (FP64, on the device)."""
        out = np.zeros(self.level(0)["nElems"])
        _check(lib().psx_element_energies(self.h, _p(np.ascontiguousarray(u, np.float64)), _p(out)))
        return out

    def stresses(self, u, E, h, principals=True):
        """Nodal von Mises (and P1, P3) of a solution u, float32 per node."""
        nN = self.level(0)["nNodes"]
        vm = np.zeros(nN, np.float32)
        p1 = np.zeros(nN, np.float32) if principals else None
        p3 = np.zeros(nN, np.float32) if principals else None
        _check(lib().psx_stresses(self.h, _p(np.ascontiguousarray(u, np.float64)), float(E), float(h), int(principals), _p(vm), _p(p1), _p(p3)))
        return vm, p1, p3

    def stress_tensors(self, u, E, h):
        out = np.zeros(6 * self.level(0)["nNodes"], np.float32)
        _check(lib().psx_stress_tensors(self.h, _p(np.ascontiguousarray(u, np.float64)), float(E), float(h), _p(out)))
        return out.reshape(-1, 6)

    def profile(self, reps=20):
        buf = C.create_string_buffer(1 << 16)
        _check(lib().psx_profile(self.h, int(reps), buf, len(buf)))
        return buf.value.decode()

    @property
    def device_bytes(self):
        return int(lib().psx_device_bytes(self.h))


# ---- HBM buffers for the *_device calls (the HIP runtime directly; no torch needed) ----

_hip = None


def _hiprt():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """A flat float64 buffer in HBM."""

    def __init__(self, n, host=None):
        self.n = int(n)
        self.ptr = C.c_void_p()
        if _hiprt().hipMalloc(C.byref(self.ptr), max(8, 8 * self.n)) != 0:
            raise PsxError(ERR_DEVICE, "hipMalloc failed")
        if host is not None:
            self.upload(host)

    def __int__(self):
        return int(self.ptr.value)

    def upload(self, host):
        h = np.ascontiguousarray(host, np.float64).reshape(-1)
        assert h.size == self.n
        if _hiprt().hipMemcpy(self.ptr, C.c_void_p(h.ctypes.data), 8 * self.n, 1) != 0:
            raise PsxError(ERR_DEVICE, "hipMemcpy H2D failed")

    def download(self):
        h = np.zeros(self.n)
        if _hiprt().hipMemcpy(C.c_void_p(h.ctypes.data), self.ptr, 8 * self.n, 2) != 0:
            raise PsxError(ERR_DEVICE, "hipMemcpy D2H failed")
        return h

    def copy_from(self, other):
        """Device-to-device copy (the runtime's copy kernel), n elements."""
        if _hiprt().hipMemcpy(self.ptr, other.ptr, 8 * min(self.n, other.n), 3) != 0:
            raise PsxError(ERR_DEVICE, "hipMemcpy D2D failed")

    def free(self):
        if self.ptr:
            _hiprt().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()


class Server:
    """A device-resident request loop (include/psx.h psx_server_*): requests are admitted as slots and
    device memory free up and run concurrently, each on its own stream."""

    def __init__(self, max_inflight=4, hbm_budget=0, pack_max=0, pack_max_dof=0):
        """pack_max: requests solved together in the pack (0 = 16, 1 = no packing); pack_max_dof: the
        largest part (DOFs) that joins it (0 = 100 000)."""
        h = C.c_void_p()
        _check(lib().psx_server_create(C.byref(ServerOpts(int(max_inflight), int(hbm_budget), int(pack_max), int(pack_max_dof))), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, dims, density, nu, bc, f, x0=None, part_id=-1, coarsest_max_dof=1100, tol=1e-6, max_iter=500, precision="fp64",
               cheb_degree=0, diag_add=None, device_build=True, mfma_smoother=0):
        """Queues one part with its load cases f (nrhs x nDof or nDof); returns a ticket.
        device_build: the request's hierarchy is built by the device kernels (False: on the host)."""
        keep = []
        d = _desc(dims, density, nu, bc, coarsest_max_dof, diag_add, 0, keep, device_build)
        f = np.ascontiguousarray(f, np.float64)
        n = 3 * int(np.prod([int(v) + 1 for v in dims]))
        nrhs = f.size // n
        if f.size != nrhs * n or nrhs < 1:
            raise ValueError("Force or displacement vector has the wrong size.")
        x0a = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        if x0a is not None and x0a.size not in (n, nrhs * n):
            raise ValueError("Force or displacement vector has the wrong size.")
        o, _ = Model._opts(tol, max_iter, precision, 0, 0, None, cheb_degree, mfma_smoother)
        req = Request(d, nrhs, f.ctypes.data, None if x0a is None else x0a.ctypes.data, int(x0a is not None and x0a.size == nrhs * n),
                      int(part_id), o)
        t = C.c_int64()
        _check(lib().psx_server_submit(self.h, C.byref(req), C.byref(t)))
        self._keep[t.value] = (nrhs, n)
        return t.value

    def wait(self, ticket):
        """Blocks until the request is done: (u as nrhs x nDof, result dict)."""
        nrhs, n = self._keep.pop(ticket)
        u = np.zeros((nrhs, n))
        r = Result()
        code = lib().psx_server_wait(self.h, int(ticket), _p(u), C.byref(r))
        out = {k: getattr(r, k) for k, _ in Result._fields_}
        if code != OK:
            raise PsxError(code, lib().psx_last_error().decode())
        return u, out

    def pending(self):
        return lib().psx_server_pending(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().psx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def solve_packed(models, d_f, d_u, nrhs, d_x0=None, tol=1e-6, max_iter=500, precision="mixed", check_every=0, fold_every=0, cheb_degree=0,
                 eager=False, mfma_smoother=0):
    """psx_solve_packed: the load cases of several models in shared launches. models: Model list;
    d_f, d_u (and d_x0, entries may be None): per model a device array of nrhs[i] x its nDof; returns
    per model the list of its load cases' info dicts (each as that model's solve_multi_device alone)."""
    n = len(models)
    o, _ = Model._opts(tol, max_iter, precision, check_every, fold_every, None, cheb_degree, mfma_smoother, eager)
    hs = (C.c_void_p * n)(*[m.h.value if isinstance(m.h, C.c_void_p) else m.h for m in models])
    ks = (C.c_int * n)(*[int(k) for k in nrhs])
    addr = lambda a: None if a is None else int(a)
    fs = (C.c_void_p * n)(*[addr(a) for a in d_f])
    us = (C.c_void_p * n)(*[addr(a) for a in d_u])
    xs = None if d_x0 is None else (C.c_void_p * n)(*[addr(a) for a in d_x0])
    infos = (SolveInfo * int(sum(ks)))()
    _check(lib().psx_solve_packed(n, hs, ks, fs, xs, C.byref(o), us, infos))
    out, at = [], 0
    for k in ks:
        out.append([Model._info(infos[at + i]) for i in range(k)])
        at += k
    return out


def prune_floating_device(dims, density, held, no_alloc=False):
    """pruneFloating on the device (psx_prune_floating_device) from host arrays: (density, removed).
    no_alloc: the allocation-free entry point (psx_prune_floating_device_noalloc, static scratch)."""
    nx, ny, nz = (int(d) for d in dims)
    density = np.ascontiguousarray(density, np.float32).reshape(-1)
    held = np.ascontiguousarray(held, np.uint8).reshape(-1)
    if density.size != nx * ny * nz or held.size != (nx + 1) * (ny + 1) * (nz + 1):
        raise ValueError("density needs one value per voxel and held one per node")
    hip = _hiprt()
    bufs = [C.c_void_p() for _ in range(3)]
    sizes = (4 * density.size, held.size, 4 * density.size)
    out = np.zeros_like(density)
    try:
        for b, n in zip(bufs, sizes):
            if hip.hipMalloc(C.byref(b), max(8, n)) != 0:
                raise PsxError(ERR_DEVICE, "hipMalloc failed")
        if hip.hipMemcpy(bufs[0], C.c_void_p(density.ctypes.data), sizes[0], 1) or hip.hipMemcpy(bufs[1], C.c_void_p(held.ctypes.data), sizes[1], 1):
            raise PsxError(ERR_DEVICE, "hipMemcpy H2D failed")
        removed = C.c_int()
        fn = lib().psx_prune_floating_device_noalloc if no_alloc else lib().psx_prune_floating_device
        _check(fn(nx, ny, nz, bufs[0], bufs[1], bufs[2], C.byref(removed)))
        if hip.hipMemcpy(C.c_void_p(out.ctypes.data), bufs[2], sizes[2], 2) != 0:
            raise PsxError(ERR_DEVICE, "hipMemcpy D2H failed")
    finally:
        for b in bufs:
            if b.value:
                hip.hipFree(b)
    return out, removed.value
