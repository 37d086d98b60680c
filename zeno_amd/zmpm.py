"""ctypes binding of libzmpm.so (include/zmpm.h) and a thin object wrapper.

Mirrors the reference's Python harness pattern (PyZpc: handles + ctypes
`CLib.call` with uint32 status, projects/PyZpc/Lib/zpy/utils.py:9-24): every C
call returns a status; a non-zero status raises ZmpmError carrying the
library's last error message.  Device memory is owned by torch tensors
(ROCm build) and handed over as raw pointers; torch is imported before the
library so both share one HIP runtime.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np
import torch  # noqa: F401  (must be loaded before libzmpm: one HIP runtime per process)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libzmpm.so")

ZMPM_OK = 0
ERRORS = {1: "ZMPM_ERR_ARG", 2: "ZMPM_ERR_HIP", 3: "ZMPM_ERR_CAPACITY", 4: "ZMPM_ERR_DOMAIN",
          5: "ZMPM_ERR_STATE", 6: "ZMPM_ERR_NODEVICE"}
MODEL_FCR = 0
MODEL_JFLUID = 1
MODEL_STVK_HENCKY = 2
MODEL_NEO_HOOKEAN = 3
MODEL_STABLE_NEO_HOOKEAN = 4
TRANSFER_APIC = 0
TRANSFER_AFLIP = 1
TRANSFER_FLIP = 2
COLLIDER_STICKY = 0
COLLIDER_SLIP = 1
COLLIDER_SEPARATE = 2
SHAPE_HALFSPACE = 0
SHAPE_SPHERE = 1
SHAPE_BOX = 2
SHAPE_SDF_GRID = 3
SHAPE_SDF_LEAVES = 4
PLASTIC_NONE = 0
PLASTIC_DRUCKER_PRAGER = 1
PLASTIC_VON_MISES = 2
PLASTIC_CAM_CLAY = 3
ANISO_NONE = 0
ANISO_ARAP = 1
BOUNDARY_NONE = 0
BOUNDARY_STICKY_BOX = 1
LAYOUT_SOA = 0
LAYOUT_AOS = 1
LAYOUT_AOSOA32 = 2

# every entry point of include/zmpm.h
EXPORTED = [
    "zmpm_create", "zmpm_destroy", "zmpm_last_error", "zmpm_set_stream", "zmpm_sync",
    "zmpm_set_materials", "zmpm_import", "zmpm_export", "zmpm_num_particles",
    "zmpm_partition", "zmpm_grid_clear", "zmpm_p2g", "zmpm_grid_update", "zmpm_boundary", "zmpm_g2p",
    "zmpm_max_vel_sqr", "zmpm_substeps", "zmpm_substeps_timed", "zmpm_num_particle_blocks", "zmpm_num_grid_blocks",
    "zmpm_num_block_parts",
    "zmpm_grid_download", "zmpm_order_download", "zmpm_bin_keys", "zmpm_sort_keys", "zmpm_key_bits",
    "zmpm_create_scenes", "zmpm_num_scenes", "zmpm_scene_max_vel_sqr", "zmpm_graph_capture", "zmpm_graph_launch",
    "zmpm_state_bytes", "zmpm_state_save", "zmpm_state_load", "zmpm_adjoint_substep", "zmpm_autotune",
    "zmpm_particle_volume", "zmpm_return_mapping", "zmpm_set_colliders", "zmpm_step_colliders",
    "zmpm_set_transfer", "zmpm_compute_particle_beta", "zmpm_set_boundary_particles",
    "zmpm_import_objects", "zmpm_export_objects", "zmpm_check_particles", "zmpm_max_vel_sqr_device",
    "zmpm_sort_keys_noalloc", "zmpm_set_fused", "zmpm_fused", "zmpm_order_by_morton_code",
]


class ZmpmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class CMaterial(ctypes.Structure):
    _fields_ = [("model", ctypes.c_int32), ("E", ctypes.c_float), ("nu", ctypes.c_float),
                ("rho", ctypes.c_float), ("vol", ctypes.c_float), ("bulk", ctypes.c_float),
                ("scene", ctypes.c_int32), ("plasticity", ctypes.c_int32), ("friction_angle", ctypes.c_float),
                ("yield_stress", ctypes.c_float), ("flip_beta", ctypes.c_float), ("nacc_beta", ctypes.c_float),
                ("nacc_xi", ctypes.c_float), ("aniso", ctypes.c_int32), ("aniso_strength", ctypes.c_float)]


class CAdjoint(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("x", ctypes.c_void_p), ("v", ctypes.c_void_p), ("C", ctypes.c_void_p),
                ("F", ctypes.c_void_p), ("mu", ctypes.c_void_p), ("lam", ctypes.c_void_p)]


class CCollider(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("shape", ctypes.c_int32), ("scene", ctypes.c_int32),
                ("dims", ctypes.c_int32 * 3), ("shape_params", ctypes.c_float * 6), ("sdf", ctypes.c_void_p),
                ("origin", ctypes.c_float * 3), ("spacing", ctypes.c_float), ("translation", ctypes.c_float * 3),
                ("translation_rate", ctypes.c_float * 3), ("scale", ctypes.c_float), ("scale_rate", ctypes.c_float),
                ("ypr_degrees", ctypes.c_float * 3), ("angular_velocity", ctypes.c_float * 3),
                ("leaf_origins", ctypes.c_void_p), ("nleaves", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def collider(type=0, shape=0, params=(0.0, 1.0, 0.0, 0.0, 0.0, 0.0), scene=0, sdf=None, origin=(0.0, 0.0, 0.0),
             spacing=1.0, translation=(0.0, 0.0, 0.0), translation_rate=(0.0, 0.0, 0.0), scale=1.0,
             scale_rate=0.0, ypr_degrees=(0.0, 0.0, 0.0), angular_velocity=(0.0, 0.0, 0.0), leaf_origins=None):
    """A zmpm_collider.  SHAPE_SDF_GRID: sdf a contiguous float32 cuda tensor (nx, ny, nz).
    SHAPE_SDF_LEAVES: sdf (nleaves, 8, 8, 8) float32 and leaf_origins (nleaves, 3) int32, both
    contiguous cuda tensors; params[0] is the background."""
    c = CCollider()
    c.type, c.shape, c.scene = int(type), int(shape), int(scene)
    p = list(params) + [0.0] * (6 - len(params))
    for k in range(6):
        c.shape_params[k] = float(p[k])
    if leaf_origins is not None:
        if sdf is None or sdf.numel() != 512 * leaf_origins.shape[0]:
            raise ValueError("SDF leaves need 512 samples per leaf")
        c.leaf_origins = ctypes.c_void_p(leaf_origins.data_ptr())
        c.nleaves = int(leaf_origins.shape[0])
        c.sdf = ctypes.c_void_p(sdf.data_ptr())
    elif sdf is not None:
        c.sdf = ctypes.c_void_p(sdf.data_ptr())
        for k in range(3):
            c.dims[k] = int(sdf.shape[k])
    for k in range(3):
        c.origin[k] = float(origin[k])
        c.translation[k] = float(translation[k])
        c.translation_rate[k] = float(translation_rate[k])
        c.ypr_degrees[k] = float(ypr_degrees[k])
        c.angular_velocity[k] = float(angular_velocity[k])
    c.spacing, c.scale, c.scale_rate = float(spacing), float(scale), float(scale_rate)
    return c


class CBoundaryParticles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("x", ctypes.c_void_p), ("v", ctypes.c_void_p), ("m", ctypes.c_void_p),
                ("nrm", ctypes.c_void_p), ("scene", ctypes.c_int32)]


class CScene(ctypes.Structure):
    _fields_ = [("dx", ctypes.c_float), ("n_cells", ctypes.c_int32), ("dt", ctypes.c_float),
                ("gravity", ctypes.c_float * 3), ("boundary", ctypes.c_int32), ("pad", ctypes.c_int32)]


class CConfig(ctypes.Structure):
    _fields_ = [("dx", ctypes.c_float), ("n_cells", ctypes.c_int32), ("capacity", ctypes.c_int64),
                ("device", ctypes.c_int32)]


class CStepParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_float), ("gravity", ctypes.c_float * 3), ("boundary", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


class CParticles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("layout", ctypes.c_int32), ("tile_channels", ctypes.c_int32),
                ("x", ctypes.c_void_p), ("x_chn", ctypes.c_int32),
                ("v", ctypes.c_void_p), ("v_chn", ctypes.c_int32),
                ("C", ctypes.c_void_p), ("C_chn", ctypes.c_int32),
                ("F", ctypes.c_void_p), ("F_chn", ctypes.c_int32),
                ("J", ctypes.c_void_p), ("J_chn", ctypes.c_int32),
                ("material", ctypes.c_void_p),
                ("beta", ctypes.c_void_p), ("beta_chn", ctypes.c_int32),
                ("logJp", ctypes.c_void_p), ("logJp_chn", ctypes.c_int32),
                ("mu", ctypes.c_void_p), ("mu_chn", ctypes.c_int32),
                ("lam", ctypes.c_void_p), ("lam_chn", ctypes.c_int32),
                ("a", ctypes.c_void_p), ("a_chn", ctypes.c_int32),
                ("m", ctypes.c_void_p), ("m_chn", ctypes.c_int32),
                ("vol", ctypes.c_void_p), ("vol_chn", ctypes.c_int32),
                ("default_material", ctypes.c_int32)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libzmpm.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m zeno_amd.build`")
    L = ctypes.CDLL(path)
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    sig = {
        "zmpm_create": (i32, [ctypes.POINTER(vp), ctypes.POINTER(CConfig)]),
        "zmpm_destroy": (i32, [vp]),
        "zmpm_last_error": (ctypes.c_char_p, []),
        "zmpm_set_stream": (i32, [vp, vp]),
        "zmpm_sync": (i32, [vp]),
        "zmpm_set_materials": (i32, [vp, ctypes.POINTER(CMaterial), i32]),
        "zmpm_import": (i32, [vp, ctypes.POINTER(CParticles)]),
        "zmpm_export": (i32, [vp, ctypes.POINTER(CParticles)]),
        "zmpm_num_particles": (i64, [vp]),
        "zmpm_partition": (i32, [vp]),
        "zmpm_grid_clear": (i32, [vp]),
        "zmpm_p2g": (i32, [vp, f32]),
        "zmpm_grid_update": (i32, [vp, f32, ctypes.POINTER(f32)]),
        "zmpm_boundary": (i32, [vp, i32, i32]),
        "zmpm_g2p": (i32, [vp, f32]),
        "zmpm_max_vel_sqr": (i32, [vp, ctypes.POINTER(f32)]),
        "zmpm_substeps": (i32, [vp, ctypes.POINTER(CStepParams), i32]),
        "zmpm_substeps_timed": (i32, [vp, ctypes.POINTER(CStepParams), i32, ctypes.POINTER(f32), i32]),
        "zmpm_num_particle_blocks": (i64, [vp]),
        "zmpm_num_grid_blocks": (i64, [vp]),
        "zmpm_num_block_parts": (i64, [vp]),
        "zmpm_grid_download": (i32, [vp, vp, vp, i64]),
        "zmpm_order_download": (i32, [vp, vp, i64]),
        "zmpm_bin_keys": (i32, [vp, vp, i64, vp]),
        "zmpm_sort_keys": (i32, [vp, vp, i64, i32, vp, vp]),
        "zmpm_key_bits": (i32, [vp]),
        "zmpm_create_scenes": (i32, [ctypes.POINTER(vp), ctypes.POINTER(CScene), i32, i64, i32]),
        "zmpm_num_scenes": (i32, [vp]),
        "zmpm_scene_max_vel_sqr": (i32, [vp, ctypes.POINTER(f32), i32]),
        "zmpm_graph_capture": (i32, [vp, ctypes.POINTER(CStepParams), i32]),
        "zmpm_graph_launch": (i32, [vp]),
        "zmpm_autotune": (i32, [vp, ctypes.POINTER(CStepParams), i32, ctypes.POINTER(i32)]),
        "zmpm_set_fused": (i32, [vp, i32]),
        "zmpm_order_by_morton_code": (i32, [vp, i32, i32, i64, vp, vp, i32, ctypes.POINTER(f32),
                                            ctypes.POINTER(f32), vp, vp, vp]),
        "zmpm_fused": (i32, [vp]),
        "zmpm_state_bytes": (i64, [vp]),
        "zmpm_state_save": (i32, [vp, vp]),
        "zmpm_state_load": (i32, [vp, vp]),
        "zmpm_adjoint_substep": (i32, [vp, ctypes.POINTER(CStepParams), ctypes.POINTER(CAdjoint)]),
        "zmpm_particle_volume": (i32, [vp, i32, i32, vp, i32, vp, i32]),
        "zmpm_return_mapping": (i32, [vp]),
        "zmpm_set_colliders": (i32, [vp, ctypes.POINTER(CCollider), i32]),
        "zmpm_step_colliders": (i32, [vp, f32]),
        "zmpm_set_transfer": (i32, [vp, i32]),
        "zmpm_compute_particle_beta": (i32, [vp, ctypes.POINTER(CStepParams), f32]),
        "zmpm_set_boundary_particles": (i32, [vp, ctypes.POINTER(CBoundaryParticles), i32]),
        "zmpm_import_objects": (i32, [vp, ctypes.POINTER(CParticles), i32]),
        "zmpm_export_objects": (i32, [vp, ctypes.POINTER(CParticles), i32]),
        "zmpm_check_particles": (i32, [ctypes.POINTER(CParticles)]),
        "zmpm_max_vel_sqr_device": (i32, [vp, ctypes.POINTER(vp)]),
        "zmpm_sort_keys_noalloc": (i32, [vp, i64, i32, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc != ZMPM_OK:
        raise ZmpmError(rc, (_lib.zmpm_last_error() or b"").decode())


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


@dataclass
class Material:
    model: int = MODEL_FCR
    E: float = 1e4
    nu: float = 0.4
    rho: float = 1000.0
    vol: float = 1e-6
    bulk: float = 0.0
    scene: int = 0
    plasticity: int = 0
    friction_angle: float = 35.0
    yield_stress: float = 1e5
    flip_beta: float = 0.0
    nacc_beta: float = 2.0   # Cam-clay "beta" (Generation.cu:80)
    nacc_xi: float = 1.0     # Cam-clay "xi" (Generation.cu:81)
    aniso: int = ANISO_NONE  # "aniso" (Generation.cu:60-63)
    aniso_strength: float = 10.0  # ARAP fiber strength (Generation.cu:62)

    @staticmethod
    def from_dict(d):
        return Material(int(d.get("model", MODEL_FCR)), float(d.get("E", 0.0)), float(d.get("nu", 0.0)),
                        float(d["rho"]), float(d["vol"]), float(d.get("bulk", 0.0)), int(d.get("scene", 0)),
                        int(d.get("plasticity", PLASTIC_NONE)), float(d.get("friction_angle", 35.0)),
                        float(d.get("yield_stress", 1e5)), float(d.get("flip_beta", 0.0)),
                        float(d.get("nacc_beta", 2.0)), float(d.get("nacc_xi", 1.0)),
                        int(d.get("aniso", ANISO_NONE)), float(d.get("aniso_strength", 10.0)))


@dataclass
class StepParams:
    dt: float
    gravity: tuple = (0.0, -9.8, 0.0)
    boundary: int = BOUNDARY_STICKY_BOX
    pad: int = 3

    def c(self):
        p = CStepParams()
        p.dt = float(self.dt)
        for k in range(3):
            p.gravity[k] = float(self.gravity[k])
        p.boundary = int(self.boundary)
        p.pad = int(self.pad)
        return p


def particle_view(n, x, v=None, C=None, F=None, J=None, material=None, layout=LAYOUT_SOA, tile_channels=0,
                  chn=None, beta=None, logJp=None, mu=None, lam=None, a=None, m=None, vol=None, default_material=0):
    """C view of caller tensors (device).  SoA: x (3, n), v (3, n), C (9, n), F (9, n), J, beta, logJp,
    mu, lam, m, vol (n,), fiber a (3, n).  For LAYOUT_AOSOA32 every tensor is the TileVector base and
    chn maps each attribute name to its first channel.  material None: every particle takes
    default_material."""
    pv = CParticles()
    pv.n = int(n)
    pv.layout = int(layout)
    pv.tile_channels = int(tile_channels)
    chn = chn or {}
    for name, t in (("x", x), ("v", v), ("C", C), ("F", F), ("J", J), ("beta", beta), ("logJp", logJp),
                    ("mu", mu), ("lam", lam), ("a", a), ("m", m), ("vol", vol)):
        setattr(pv, name, t.data_ptr() if t is not None else None)
        setattr(pv, name + "_chn", int(chn.get(name, 0)))
    pv.material = material.data_ptr() if material is not None else None
    pv.default_material = int(default_material)
    return pv


def check_particles(view: CParticles):
    """Host-side check of a view (zmpm_check_particles); raises ZmpmError(ZMPM_ERR_ARG) if it fails."""
    L = load_library()
    _check(L.zmpm_check_particles(ctypes.byref(view)))


class Context:
    """An MPM context on one GPU: resident particles, partition and grid (zmpm_ctx).

    Context(dx, n_cells, capacity) holds one scene; Context.scenes([...], capacity)
    holds several independent scenes that every substep launch covers together.
    """

    def __init__(self, dx: float, n_cells: int, capacity: int, device: int = 0, _handle=None):
        L = load_library()
        if _handle is None:
            cfg = CConfig(float(dx), int(n_cells), int(capacity), int(device))
            h = ctypes.c_void_p()
            _check(L.zmpm_create(ctypes.byref(h), ctypes.byref(cfg)))
        else:
            h = _handle
        self._h = h
        self.L = L
        self.dx = float(dx)
        self.n_cells = int(n_cells)
        self.capacity = int(capacity)
        self.device = int(device)

    @classmethod
    def scenes(cls, scenes, capacity: int, device: int = 0):
        """scenes: dicts with dx, n_cells and optionally dt, gravity, boundary, pad."""
        L = load_library()
        arr = (CScene * len(scenes))()
        for i, sc in enumerate(scenes):
            g = sc.get("gravity", (0.0, -9.8, 0.0))
            arr[i] = CScene(float(sc["dx"]), int(sc["n_cells"]), float(sc.get("dt", 0.0)),
                            (ctypes.c_float * 3)(*[float(v) for v in g]), int(sc.get("boundary", BOUNDARY_STICKY_BOX)),
                            int(sc.get("pad", 3)))
        h = ctypes.c_void_p()
        _check(L.zmpm_create_scenes(ctypes.byref(h), arr, len(scenes), int(capacity), int(device)))
        return cls(float(scenes[0]["dx"]), int(scenes[0]["n_cells"]), capacity, device, _handle=h)

    def num_scenes(self):
        return int(self.L.zmpm_num_scenes(self._h))

    def scene_max_vel_sqr(self):
        n = self.num_scenes()
        out = (ctypes.c_float * n)()
        _check(self.L.zmpm_scene_max_vel_sqr(self._h, out, n))
        return np.array(out[:], np.float32)

    def graph_capture(self, params, nsteps: int):
        """Capture nsteps (even) substeps into a HIP graph; params None = each scene's own."""
        p = params.c() if params is not None else None
        _check(self.L.zmpm_graph_capture(self._h, ctypes.byref(p) if p is not None else None, int(nsteps)))

    def graph_launch(self):
        _check(self.L.zmpm_graph_launch(self._h))

    def order_by_morton_code(self, layout, nch, n, src, dst, x_chn, bv_min, bv_max, codes, ordered_to_original,
                             original_to_ordered):
        """ZenoParticles::orderByMortonCode on a caller buffer (device tensors): codes (optional,
        int32/uint32 (n,)) receives the sorted codes, the maps are int32 (n,), dst the permuted channels."""
        lo = (ctypes.c_float * 3)(*[float(v) for v in bv_min])
        hi = (ctypes.c_float * 3)(*[float(v) for v in bv_max])
        _check(self.L.zmpm_order_by_morton_code(self._h, int(layout), int(nch), int(n), _ptr(src), _ptr(dst),
                                                int(x_chn), lo, hi, _ptr(codes) if codes is not None else None,
                                                _ptr(ordered_to_original), _ptr(original_to_ordered)))

    def set_fused(self, on: bool):
        """The fused G2P / P2G pass on or off (APIC); results are bit-identical either way."""
        _check(self.L.zmpm_set_fused(self._h, 1 if on else 0))

    def fused(self) -> bool:
        return self.L.zmpm_fused(self._h) == 1

    def autotune(self, params, nsteps: int = 4):
        """Pick the G2P workgroup sizes, whether sparse particle blocks take the one-wave P2G and
        whether the fused G2P / P2G pass runs, by timing; the state is restored.
        Returns (move, update, sparse_p2g); update is 0 when the fused pass won."""
        out = (ctypes.c_int32 * 3)()
        p = ctypes.byref(params.c()) if params is not None else None
        _check(self.L.zmpm_autotune(self._h, p, int(nsteps), out))
        return int(out[0]), int(out[1]), bool(out[2])

    # -- differentiation
    def state_bytes(self):
        return int(self.L.zmpm_state_bytes(self._h))

    def new_state_buffer(self):
        return torch.empty(self.state_bytes() // 4, dtype=torch.float32, device="cuda:%d" % self.device)

    def state_save(self, buf):
        _check(self.L.zmpm_state_save(self._h, ctypes.c_void_p(buf.data_ptr())))

    def state_load(self, buf):
        _check(self.L.zmpm_state_load(self._h, ctypes.c_void_p(buf.data_ptr())))

    def adjoint_substep(self, params, adj: dict):
        """adj: device tensors x (3,n), v (3,n), C (9,n), F (9,n), mu (n,), lam (n,), by caller index;
        the substep's output adjoints on entry, its input adjoints on exit (mu, lam accumulate)."""
        a = CAdjoint(self.num_particles(), *[adj[k].data_ptr() for k in ("x", "v", "C", "F", "mu", "lam")])
        p = ctypes.byref(params.c()) if params is not None else None
        _check(self.L.zmpm_adjoint_substep(self._h, p, ctypes.byref(a)))

    def close(self):
        if getattr(self, "_h", None):
            self.L.zmpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- setup
    def set_stream(self, stream):
        _check(self.L.zmpm_set_stream(self._h, ctypes.c_void_p(stream)))

    def set_materials(self, mats):
        arr = (CMaterial * len(mats))()
        for i, m in enumerate(mats):
            m = m if isinstance(m, Material) else Material.from_dict(m)
            arr[i] = CMaterial(m.model, m.E, m.nu, m.rho, m.vol, m.bulk, m.scene, m.plasticity, m.friction_angle,
                               m.yield_stress, m.flip_beta, m.nacc_beta, m.nacc_xi, m.aniso, m.aniso_strength)
        _check(self.L.zmpm_set_materials(self._h, arr, len(mats)))

    def import_particles(self, view: CParticles):
        _check(self.L.zmpm_import(self._h, ctypes.byref(view)))

    def export_particles(self, view: CParticles):
        _check(self.L.zmpm_export(self._h, ctypes.byref(view)))

    def import_objects(self, views):
        """A list of particle objects (one CParticles view each), read in place (zmpm_import_objects)."""
        arr = (CParticles * len(views))(*views)
        _check(self.L.zmpm_import_objects(self._h, arr, len(views)))

    def export_objects(self, views):
        arr = (CParticles * len(views))(*views)
        _check(self.L.zmpm_export_objects(self._h, arr, len(views)))

    def max_vel_sqr_device(self):
        """Device address of the per-scene MaxVelSqr every fused substep writes (float32 each)."""
        p = ctypes.c_void_p()
        _check(self.L.zmpm_max_vel_sqr_device(self._h, ctypes.byref(p)))
        return int(p.value or 0)

    def import_state(self, state: dict):
        """state: device tensors x (3,n), v (3,n), C (9,n), F (9,n) and mat (n,) int32; J taken from F[0];
        optional beta (n,) and logJp (n,) (default: the material's flip_beta, and -0.04), and per-particle
        Lame parameters mu, lam (n,) (both or neither; default: the material's) and the ARAP fiber a (3, n)
        (default (0, 1, 0)), and per-particle mass and volume m, vol (n,) (both or neither; default: the
        material's rho * vol and vol)."""
        n = state["x"].shape[1]
        self._keep = state
        J = state["F"][0].contiguous()
        self._keepJ = J
        self.import_particles(particle_view(n, state["x"], state["v"], state["C"], state["F"], J, state["mat"],
                                            beta=state.get("beta"), logJp=state.get("logJp"), mu=state.get("mu"),
                                            lam=state.get("lam"), a=state.get("a"), m=state.get("m"),
                                            vol=state.get("vol")))

    def export_state(self, device="cuda", aux=False):
        """x, v, C, F (J-fluid: J in F[0]); with aux also beta, logJp, mu, lam, m and vol (n,) and a (3, n)."""
        n = self.num_particles()
        out = {k: torch.empty((c, n), dtype=torch.float32, device=device) for k, c in
               (("x", 3), ("v", 3), ("C", 9), ("F", 9))}
        J = torch.empty(n, dtype=torch.float32, device=device)
        out["F"].zero_()
        J.fill_(float("nan"))
        if aux:
            for k in ("beta", "logJp", "mu", "lam", "m", "vol"):
                out[k] = torch.empty(n, dtype=torch.float32, device=device)
            out["a"] = torch.empty((3, n), dtype=torch.float32, device=device)
        self.export_particles(particle_view(n, out["x"], out["v"], out["C"], out["F"], J, None,
                                            beta=out.get("beta"), logJp=out.get("logJp"), mu=out.get("mu"),
                                            lam=out.get("lam"), a=out.get("a"), m=out.get("m"), vol=out.get("vol")))
        # the oracle convention: J-fluid particles carry J in F[0] (J is NaN for F-based particles)
        fluid = ~torch.isnan(J)
        out["F"][0] = torch.where(fluid, J, out["F"][0])
        return out

    def num_particles(self):
        return int(self.L.zmpm_num_particles(self._h))

    def particle_volume(self, device="cuda"):
        """ComputeParticleVolume (Maintenance.cu:17-56): per-particle dx^3 / (particles in the
        cell floor(x / dx)) and rho * vol, in the caller's particle order. Returns (vol, mass).
        The resident particles keep them as their own volume and mass from then on, which P2G
        reads (Transfer.cu:857-858)."""
        n = self.num_particles()
        vol = torch.empty(n, dtype=torch.float32, device=device)
        mass = torch.empty(n, dtype=torch.float32, device=device)
        _check(self.L.zmpm_particle_volume(self._h, LAYOUT_SOA, 0, ctypes.c_void_p(vol.data_ptr()), 0,
                                           ctypes.c_void_p(mass.data_ptr()), 0))
        self.sync()
        return vol, mass

    def set_colliders(self, cols, keep=()):
        """Level-set colliders (ToZSBoundary): a list of CCollider (see collider()); keep: tensors
        (SDF samples) that must outlive the colliders."""
        arr = (CCollider * max(len(cols), 1))(*cols)
        self._keep_cols = (arr, tuple(keep))
        _check(self.L.zmpm_set_colliders(self._h, arr, len(cols)))

    def set_transfer(self, scheme):
        """TRANSFER_APIC, TRANSFER_AFLIP (ASFLIP through the particles' beta) or TRANSFER_FLIP."""
        _check(self.L.zmpm_set_transfer(self._h, int(scheme)))

    def compute_particle_beta(self, params=None, jp_critical=None):
        """ComputeParticleBeta with the colliders as the boundary (Boundary.cu:284-313): beta 0 for
        particles moving into a collider at x + dt v, else 0.5 where det F > jp_critical, else 0.
        jp_critical None: the reference's JpcDefault, 1 for every material (its 1.1 for Cam-clay
        tests a model variant that cannot hold Cam-clay, Boundary.cu:389-391)."""
        jp_critical = -1.0 if jp_critical is None else jp_critical
        p = ctypes.byref(params.c()) if params is not None else None
        _check(self.L.zmpm_compute_particle_beta(self._h, p, float(jp_critical)))

    def set_boundary_particles(self, x=None, v=None, m=None, nrm=None, type=COLLIDER_STICKY, scene=0):
        """Boundary particles (ZSBoundaryPrimitiveToZSGrid + ApplyGridBoundaryOnZSGrid): contiguous
        float32 cuda tensors x, v, nrm (3, n) and m (n,), read at every substep (kept alive here);
        x None removes the boundary."""
        if x is None:
            self._keep_bnd = None
            _check(self.L.zmpm_set_boundary_particles(self._h, None, int(type)))
            return
        ts = (x, v, m, nrm)
        self._keep_bnd = ts
        bp = CBoundaryParticles(int(x.shape[1]), *[ctypes.c_void_p(t.data_ptr()) if t is not None else None
                                                   for t in ts], int(scene))
        _check(self.L.zmpm_set_boundary_particles(self._h, ctypes.byref(bp), int(type)))

    def step_colliders(self, dt):
        """StepZSBoundary: translation += rate dt, scale += rate dt."""
        _check(self.L.zmpm_step_colliders(self._h, float(dt)))

    def return_mapping(self):
        """ZSReturnMapping (MPMPipeline.cu:953-972): project F of the plastic particles."""
        _check(self.L.zmpm_return_mapping(self._h))

    def particle_volume_into(self, layout, tile_channels, vol, vol_chn, mass, mass_chn):
        """ComputeParticleVolume into caller buffers: one channel each of a LAYOUT_* buffer
        (for LAYOUT_AOSOA32 the TileVector base and its "vol" / "m" channels)."""
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        _check(self.L.zmpm_particle_volume(self._h, int(layout), int(tile_channels), ptr(vol), int(vol_chn),
                                           ptr(mass), int(mass_chn)))

    # -- node chain
    def partition(self):
        _check(self.L.zmpm_partition(self._h))

    def grid_clear(self):
        _check(self.L.zmpm_grid_clear(self._h))

    def p2g(self, dt):
        _check(self.L.zmpm_p2g(self._h, float(dt)))

    def grid_update(self, dt, accel=(0.0, -9.8, 0.0)):
        a = (ctypes.c_float * 3)(*[float(v) for v in accel])
        _check(self.L.zmpm_grid_update(self._h, float(dt), a))

    def boundary(self, kind=BOUNDARY_STICKY_BOX, pad=3):
        _check(self.L.zmpm_boundary(self._h, int(kind), int(pad)))

    def g2p(self, dt):
        _check(self.L.zmpm_g2p(self._h, float(dt)))

    def max_vel_sqr(self):
        v = ctypes.c_float()
        _check(self.L.zmpm_max_vel_sqr(self._h, ctypes.byref(v)))
        return v.value

    def substeps(self, params, nsteps: int):
        """nsteps fused substeps; params None runs every scene with its own parameters."""
        if params is None:
            _check(self.L.zmpm_substeps(self._h, None, int(nsteps)))
        else:
            p = params.c()
            _check(self.L.zmpm_substeps(self._h, ctypes.byref(p), int(nsteps)))

    PHASES = ("stress", "p2g", "grid", "g2p_move", "partition", "g2p_update")

    def substeps_timed(self, params: StepParams, nsteps: int):
        """Run nsteps substeps with per-phase HIP events; returns {phase: summed GPU ms}."""
        out = (ctypes.c_float * len(self.PHASES))()
        p = ctypes.byref(params.c()) if params is not None else None
        _check(self.L.zmpm_substeps_timed(self._h, p, int(nsteps), out, len(self.PHASES)))
        return dict(zip(self.PHASES, [float(v) for v in out]))

    def sync(self):
        _check(self.L.zmpm_sync(self._h))

    # -- introspection
    def num_particle_blocks(self):
        return int(self.L.zmpm_num_particle_blocks(self._h))

    def num_grid_blocks(self):
        return int(self.L.zmpm_num_grid_blocks(self._h))

    def num_block_parts(self):
        """Work units of the current partition: blocks of more than 1,024 particles count several."""
        return int(self.L.zmpm_num_block_parts(self._h))

    def grid(self):
        G = self.num_grid_blocks()
        keys = np.empty((max(G, 1), 3), np.int32)
        data = np.empty((max(G, 1), 4, 64), np.float32)
        _check(self.L.zmpm_grid_download(self._h, keys.ctypes.data_as(ctypes.c_void_p),
                                         data.ctypes.data_as(ctypes.c_void_p), G))
        return keys[:G], data[:G]

    def order(self):
        n = self.num_particles()
        o = np.empty(n, np.int32)
        _check(self.L.zmpm_order_download(self._h, o.ctypes.data_as(ctypes.c_void_p), n))
        return o

    def key_bits(self):
        return int(self.L.zmpm_key_bits(self._h))

    def bin_keys(self, x):
        n = x.shape[1]
        keys = torch.empty(n, dtype=torch.int32, device=x.device)
        _check(self.L.zmpm_bin_keys(self._h, _ptr(x.contiguous()), n, _ptr(keys)))
        return keys

    def sort_keys(self, keys, key_bits):
        n = keys.numel()
        ko = torch.empty_like(keys)
        po = torch.empty_like(keys)
        _check(self.L.zmpm_sort_keys(self._h, _ptr(keys), n, int(key_bits), _ptr(ko), _ptr(po)))
        return ko, po
