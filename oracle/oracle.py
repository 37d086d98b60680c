"""ctypes binding of the CPU oracle (oracle/liboracle_mpm.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product (zeno_amd / libzmpm.so)
never imports it.

Particle state is SoA: x, v are (3, n); C, F are (9, n) with entry (r, c) of
the 3x3 matrix in row 3*r+c; mat is (n,) int32 indices into the material
table.  For J-fluid particles F[0] holds J and F[1:] is unused.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_mpm.so")

MODEL_FCR = 0
MODEL_JFLUID = 1
MODEL_STVK_HENCKY = 2
MODEL_NEO_HOOKEAN = 3
MODEL_STABLE_NEO_HOOKEAN = 4
TRANSFER_APIC = 0
TRANSFER_AFLIP = 1
TRANSFER_FLIP = 2
PLASTIC_NONE = 0
PLASTIC_DRUCKER_PRAGER = 1
PLASTIC_VON_MISES = 2
PLASTIC_CAM_CLAY = 3
ANISO_NONE = 0
ANISO_ARAP = 1
BOUNDARY_NONE = 0
BOUNDARY_STICKY_BOX = 1


class Params(ctypes.Structure):
    _fields_ = [
        ("dx", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("n_cells", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("threads", ctypes.c_int32),
    ]


class Material(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32),
        ("E", ctypes.c_float),
        ("nu", ctypes.c_float),
        ("rho", ctypes.c_float),
        ("vol", ctypes.c_float),
        ("bulk", ctypes.c_float),
        ("plasticity", ctypes.c_int32),
        ("friction_angle", ctypes.c_float),
        ("yield_stress", ctypes.c_float),
        ("flip_beta", ctypes.c_float),
        ("nacc_beta", ctypes.c_float),
        ("nacc_xi", ctypes.c_float),
        ("aniso", ctypes.c_int32),
        ("aniso_strength", ctypes.c_float),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        for sfx in ("f32", "f64"):
            getattr(L, f"orc_create_{sfx}").restype = vp
            getattr(L, f"orc_create_{sfx}").argtypes = [vp, vp, i32, i64, vp, vp, vp, vp, vp]
            getattr(L, f"orc_destroy_{sfx}").argtypes = [vp]
            getattr(L, f"orc_set_dt_{sfx}").argtypes = [vp, f64]
            getattr(L, f"orc_set_threads_{sfx}").argtypes = [vp, i32]
            getattr(L, f"orc_stage_{sfx}").argtypes = [vp, i32]
            getattr(L, f"orc_stage_{sfx}").restype = i32
            getattr(L, f"orc_substeps_{sfx}").argtypes = [vp, i32]
            getattr(L, f"orc_max_vel_sqr_{sfx}").argtypes = [vp]
            getattr(L, f"orc_max_vel_sqr_{sfx}").restype = f64
            getattr(L, f"orc_get_{sfx}").argtypes = [vp, vp, vp, vp, vp]
            getattr(L, f"orc_num_gblocks_{sfx}").argtypes = [vp]
            getattr(L, f"orc_num_gblocks_{sfx}").restype = i64
            getattr(L, f"orc_grid_{sfx}").argtypes = [vp, vp, vp]
        L.orc_num_pblocks_f32.argtypes = [vp]
        L.orc_num_pblocks_f32.restype = i64
        L.orc_porder_f32.argtypes = [vp, vp]
        L.orc_svd3.argtypes = [vp, vp, vp, vp]
        L.orc_polar_rotation.argtypes = [vp, vp]
        L.orc_fcr_piola_f64.argtypes = [vp, f64, f64, vp]
        L.orc_fcr_piola_f32.argtypes = [vp, ctypes.c_float, ctypes.c_float, vp]
        L.orc_lame_f32.argtypes = [ctypes.c_float, ctypes.c_float, vp, vp]
        L.orc_neo_hookean_piola_f64.argtypes = [vp, f64, f64, i32, vp]
        L.orc_arap_fiber_piola_f64.argtypes = [vp, vp, f64, vp]
        L.orc_stvk_hencky_tau_f64.argtypes = [vp, f64, f64, vp]
        L.orc_set_colliders_f32.argtypes = [vp, vp, i32]
        L.orc_set_transfer_f32.argtypes = [vp, i32]
        L.orc_set_transfer_f64.argtypes = [vp, i32]
        L.orc_set_colliders_f64.argtypes = [vp, vp, i32]
        L.orc_resolve_collision_f64.argtypes = [vp, vp, vp]
        L.orc_resolve_collision_f64.restype = i32
        L.orc_shape_phi_f64.argtypes = [vp, vp, vp]
        L.orc_shape_phi_f64.restype = ctypes.c_double
        L.orc_project_strain_f64.argtypes = [vp, f64, f64, i32, f64, f64]
        L.orc_project_strain_f64.restype = i32
        L.orc_arena_f64.argtypes = [vp, f64, vp, vp, vp]
        L.orc_weight_grad_f64.argtypes = [vp, f64, vp, vp, vp]
        L.orc_gstar_f64.argtypes = [vp, vp]
        L.orc_project_cam_clay_f64.argtypes = [vp, f64, f64, f64, f64, f64, vp]
        L.orc_project_cam_clay_f64.restype = i32
        for sfx in ("f32", "f64"):
            getattr(L, f"orc_aux_{sfx}").argtypes = [vp, vp, vp, i32]
            getattr(L, f"orc_compute_particle_beta_{sfx}").argtypes = [vp, f64, f64]
            getattr(L, f"orc_set_boundary_particles_{sfx}").argtypes = [vp, i64, vp, vp, vp, vp, i32]
            getattr(L, f"orc_set_lame_{sfx}").argtypes = [vp, vp, vp]
            getattr(L, f"orc_set_mass_vol_{sfx}").argtypes = [vp, vp, vp]
            getattr(L, f"orc_set_fiber_{sfx}").argtypes = [vp, vp]
        L.orc_cell_coords_f32.argtypes = [vp, i64, ctypes.c_float, vp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def make_params(dx, dt, gravity=(0.0, -9.8, 0.0), n_cells=None, pad=3, boundary=BOUNDARY_STICKY_BOX, threads=0):
    p = Params()
    p.dx = float(dx)
    p.dt = float(dt)
    for k in range(3):
        p.gravity[k] = float(gravity[k])
    p.n_cells = int(round(1.0 / dx)) if n_cells is None else int(n_cells)
    p.pad = int(pad)
    p.boundary = int(boundary)
    p.threads = int(threads) if threads else (os.cpu_count() or 1)
    return p


def make_materials(mats):
    """mats: sequence of dicts with model, E, nu, rho, vol, bulk."""
    arr = (Material * len(mats))()
    for i, m in enumerate(mats):
        arr[i].model = int(m.get("model", MODEL_FCR))
        arr[i].E = float(m.get("E", 0.0))
        arr[i].nu = float(m.get("nu", 0.0))
        arr[i].rho = float(m["rho"])
        arr[i].vol = float(m["vol"])
        arr[i].bulk = float(m.get("bulk", 0.0))
        arr[i].plasticity = int(m.get("plasticity", PLASTIC_NONE))
        arr[i].friction_angle = float(m.get("friction_angle", 35.0))
        arr[i].yield_stress = float(m.get("yield_stress", 1e5))
        arr[i].flip_beta = float(m.get("flip_beta", 0.0))
        arr[i].nacc_beta = float(m.get("nacc_beta", 2.0))
        arr[i].nacc_xi = float(m.get("nacc_xi", 1.0))
        arr[i].aniso = int(m.get("aniso", ANISO_NONE))
        arr[i].aniso_strength = float(m.get("aniso_strength", 10.0))
    return arr


class Session:
    """One oracle simulation state (FP32 or FP64), stepped by stage like the zeno node chain."""

    STAGES = {"partition": 0, "p2g": 1, "grid_update": 2, "boundary": 3, "g2p": 4}

    def __init__(self, state: dict, mats, dx, dt, gravity=(0.0, -9.8, 0.0), pad=3,
                 boundary=BOUNDARY_STICKY_BOX, n_cells=None, dtype=np.float32, threads=0):
        self.dtype = np.dtype(dtype)
        self.sfx = "f32" if self.dtype == np.float32 else "f64"
        L = lib()
        self._params = make_params(dx, dt, gravity, n_cells, pad, boundary, threads)
        self._mats = make_materials(mats)
        n = int(state["x"].shape[1])
        self.n = n
        cast = lambda a, k: np.ascontiguousarray(np.asarray(a, dtype=self.dtype).reshape(k, n))
        x, v, C, F = cast(state["x"], 3), cast(state["v"], 3), cast(state["C"], 9), cast(state["F"], 9)
        mat = np.ascontiguousarray(np.asarray(state["mat"], dtype=np.int32).reshape(n))
        self._h = getattr(L, f"orc_create_{self.sfx}")(
            ctypes.byref(self._params), ctypes.byref(self._mats), len(mats), n,
            _ptr(x), _ptr(v), _ptr(C), _ptr(F), _ptr(mat))

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            getattr(_lib, f"orc_destroy_{self.sfx}")(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_transfer(self, scheme):
        """0: APIC, 1: AFLIP/ASFLIP (the particles' beta), 2: FLIP without the affine term."""
        getattr(lib(), f"orc_set_transfer_{self.sfx}")(self._h, int(scheme))

    def set_aux(self, beta=None, logJp=None):
        """Per-particle beta and logJp (n,) (initially the material's flip_beta and -0.04)."""
        b = None if beta is None else np.ascontiguousarray(beta, dtype=self.dtype).reshape(self.n)
        lj = None if logJp is None else np.ascontiguousarray(logJp, dtype=self.dtype).reshape(self.n)
        getattr(lib(), f"orc_aux_{self.sfx}")(self._h, None if b is None else _ptr(b),
                                               None if lj is None else _ptr(lj), 1)

    def aux(self):
        """(beta, logJp) per particle."""
        b = np.empty(self.n, self.dtype)
        lj = np.empty(self.n, self.dtype)
        getattr(lib(), f"orc_aux_{self.sfx}")(self._h, _ptr(b), _ptr(lj), 0)
        return b, lj

    def set_boundary_particles(self, x, v, m, nrm=None, type=0):
        """Boundary particles x, v, nrm (3, nb) and m (nb,) (float32 values), applied after the
        colliders with response type 0 sticky, 1 slip, 2 separate."""
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        self._bnd = [f(x), f(v), f(m), f(nrm)]
        nb = self._bnd[0].shape[1]
        getattr(lib(), f"orc_set_boundary_particles_{self.sfx}")(
            self._h, nb, *[None if a is None else _ptr(a) for a in self._bnd], int(type))

    def set_lame(self, mu=None, lam=None):
        """Per-particle Lame parameters (n,) each, in place of the materials' (Transfer.cu:861-864);
        None removes them."""
        if mu is None or lam is None:
            self._lame = None
            getattr(lib(), f"orc_set_lame_{self.sfx}")(self._h, None, None)
            return
        self._lame = [np.ascontiguousarray(mu, dtype=np.float32), np.ascontiguousarray(lam, dtype=np.float32)]
        getattr(lib(), f"orc_set_lame_{self.sfx}")(self._h, _ptr(self._lame[0]), _ptr(self._lame[1]))

    def set_mass_vol(self, m=None, vol=None):
        """Per-particle mass and volume (n,) each, in place of the materials' rho * vol and vol
        (pars("m"), pars("vol"), Transfer.cu:857-858); None removes them."""
        if m is None or vol is None:
            self._mv = None
            getattr(lib(), f"orc_set_mass_vol_{self.sfx}")(self._h, None, None)
            return
        self._mv = [np.ascontiguousarray(m, dtype=np.float32), np.ascontiguousarray(vol, dtype=np.float32)]
        getattr(lib(), f"orc_set_mass_vol_{self.sfx}")(self._h, _ptr(self._mv[0]), _ptr(self._mv[1]))

    def set_fiber(self, a=None):
        """Per-particle rest-space fiber directions (3, n) of the ARAP fiber term (the "a"
        channel, Generation.cpp:892,977-982); None: (0, 1, 0) for every particle."""
        if a is None:
            self._fiber = None
            getattr(lib(), f"orc_set_fiber_{self.sfx}")(self._h, None)
            return
        self._fiber = np.ascontiguousarray(a, dtype=np.float32).reshape(3, -1)
        getattr(lib(), f"orc_set_fiber_{self.sfx}")(self._h, _ptr(self._fiber))

    def compute_particle_beta(self, dt, jp_critical=None):
        """ComputeParticleBeta with the session's colliders as the level-set boundary; jp_critical
        None: the reference's JpcDefault, 1 for every material."""
        jp_critical = -1.0 if jp_critical is None else jp_critical
        getattr(lib(), f"orc_compute_particle_beta_{self.sfx}")(self._h, float(dt), float(jp_critical))

    def set_colliders(self, cols):
        """cols: dicts with the zmpm_collider fields (make_collider)."""
        self._col_keep = []
        arr = (Collider * max(len(cols), 1))(*[make_collider(d, self._col_keep) for d in cols])
        self._col_arr = arr
        getattr(lib(), f"orc_set_colliders_{self.sfx}")(self._h, arr, len(cols))

    def set_dt(self, dt):
        getattr(lib(), f"orc_set_dt_{self.sfx}")(self._h, float(dt))

    def set_threads(self, t):
        getattr(lib(), f"orc_set_threads_{self.sfx}")(self._h, int(t))

    def stage(self, name):
        rc = getattr(lib(), f"orc_stage_{self.sfx}")(self._h, self.STAGES[name])
        assert rc == 0

    def substeps(self, k):
        getattr(lib(), f"orc_substeps_{self.sfx}")(self._h, int(k))

    def max_vel_sqr(self):
        return getattr(lib(), f"orc_max_vel_sqr_{self.sfx}")(self._h)

    def state(self):
        n = self.n
        x = np.empty((3, n), self.dtype)
        v = np.empty((3, n), self.dtype)
        C = np.empty((9, n), self.dtype)
        F = np.empty((9, n), self.dtype)
        getattr(lib(), f"orc_get_{self.sfx}")(self._h, _ptr(x), _ptr(v), _ptr(C), _ptr(F))
        return {"x": x, "v": v, "C": C, "F": F}

    def grid(self):
        """Active grid blocks: (keys (G,3) int32 block origin in cells, data (G,4,64)); cell id = (lx*4+ly)*4+lz."""
        G = getattr(lib(), f"orc_num_gblocks_{self.sfx}")(self._h)
        keys = np.empty((G, 3), np.int32)
        data = np.empty((G, 4, 64), self.dtype)
        getattr(lib(), f"orc_grid_{self.sfx}")(self._h, _ptr(keys), _ptr(data))
        return keys, data

    def gstar(self):
        """FP64 sessions with a FLIP transfer: the vstar grid (G, 4, 64), blocks as in grid()."""
        assert self.sfx == "f64"
        G = lib().orc_num_gblocks_f64(self._h)
        data = np.empty((G, 4, 64), np.float64)
        lib().orc_gstar_f64(self._h, _ptr(data))
        return data

    def porder(self):
        assert self.sfx == "f32"
        o = np.empty(self.n, np.int64)
        lib().orc_porder_f32(self._h, _ptr(o))
        return o


# ---- primitives
def svd3(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    U = np.empty(9)
    s = np.empty(3)
    V = np.empty(9)
    lib().orc_svd3(_ptr(F), _ptr(U), _ptr(s), _ptr(V))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def polar_rotation(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    R = np.empty(9)
    lib().orc_polar_rotation(_ptr(F), _ptr(R))
    return R.reshape(3, 3)


def fcr_piola(F, mu, lam, dtype=np.float64):
    if np.dtype(dtype) == np.float64:
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
        P = np.empty(9)
        lib().orc_fcr_piola_f64(_ptr(F), float(mu), float(lam), _ptr(P))
    else:
        F = np.ascontiguousarray(F, dtype=np.float32).reshape(9)
        P = np.empty(9, np.float32)
        lib().orc_fcr_piola_f32(_ptr(F), float(mu), float(lam), _ptr(P))
    return P.reshape(3, 3)


def neo_hookean_piola(F, mu, lam, stable=False):
    """First Piola stress of the Neo-Hookean (stable=False) or stable Neo-Hookean model (FP64)."""
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    P = np.empty(9)
    lib().orc_neo_hookean_piola_f64(_ptr(F), float(mu), float(lam), int(bool(stable)), _ptr(P))
    return P.reshape(3, 3)


def arap_fiber_piola(F, a, mu_a):
    """The anisotropic ARAP fiber term of P for fiber a and stiffness mu_a (FP64)."""
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(3)
    P = np.empty(9)
    lib().orc_arap_fiber_piola_f64(_ptr(F), _ptr(a), float(mu_a), _ptr(P))
    return P.reshape(3, 3)


def lame_f32(E, nu):
    mu = ctypes.c_float()
    lam = ctypes.c_float()
    lib().orc_lame_f32(float(E), float(nu), ctypes.byref(mu), ctypes.byref(lam))
    return mu.value, lam.value


def arena(xp, dx):
    xp = np.ascontiguousarray(xp, dtype=np.float64).reshape(3)
    base = np.empty(3, np.int32)
    w = np.empty(9)
    d = np.empty(9)
    lib().orc_arena_f64(_ptr(xp), float(dx), _ptr(base), _ptr(w), _ptr(d))
    return base, w.reshape(3, 3), d.reshape(3, 3)


def weight_grad(xp, dx):
    """Stencil weights W (3,3,3) and gradients grad W (3,3,3,3) (d/dx_p) of one particle, FP64."""
    xp = np.ascontiguousarray(xp, dtype=np.float64).reshape(3)
    base = np.empty(3, np.int32)
    W = np.empty(27)
    g = np.empty(81)
    lib().orc_weight_grad_f64(_ptr(xp), float(dx), _ptr(base), _ptr(W), _ptr(g))
    return base, W.reshape(3, 3, 3), g.reshape(3, 3, 3, 3)


def cell_coords(x, dx):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[1]
    out = np.empty((3, n), np.int32)
    lib().orc_cell_coords_f32(_ptr(x), n, float(dx), _ptr(out))
    return out


def particle_volume(x, dx, rho, group=None):
    """ComputeParticleVolume (projects/CuLagrange/mpm/Maintenance.cu:17-56) with the
    per-cell counts of spatial_hashing (projects/CUDA/Utils.hpp:770-843), in float32:
    bucket = floor(x * (1/dx)) per axis, vol = (dx dx dx) / count(bucket),
    m = rho * vol (Maintenance.cu:41-52).  zpc's IndexBuckets::bucketCoord is absent;
    the bucket is taken as the cell containing x (one bucket per grid cell).

    x: (3, n) float32; dx: scalar or (n,) per particle; rho: scalar or (n,);
    group: optional (n,) ints, particles of different groups (scenes) never share a
    bucket.  Returns (vol, mass) as float32 arrays of n."""
    x = np.asarray(x, np.float32)
    n = x.shape[1]
    dx32 = np.broadcast_to(np.asarray(dx, np.float32), (n,))
    dxinv = np.float32(1.0) / dx32
    cells = np.floor(x * dxinv[None, :]).astype(np.int64)
    g = np.zeros(n, np.int64) if group is None else np.asarray(group, np.int64)
    key = np.concatenate([g[None, :], cells], axis=0).T
    _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    vol = ((dx32 * dx32) * dx32) / cnt[inv].astype(np.float32)
    mass = np.broadcast_to(np.asarray(rho, np.float32), (n,)) * vol
    return vol.astype(np.float32), mass.astype(np.float32)


def stvk_hencky_tau(F, mu, lam):
    """Kirchhoff stress of StVK with Hencky strain (FP64, signed SVD); F: (3, 3)."""
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    out = np.empty(9)
    lib().orc_stvk_hencky_tau_f64(_ptr(F), float(mu), float(lam), _ptr(out))
    return out.reshape(3, 3)


def project_strain(F, mu, lam, plasticity, friction_angle=35.0, yield_stress=1e5):
    """Return mapping of one deformation gradient (FP64). Returns (F', changed)."""
    F = np.array(F, dtype=np.float64).reshape(9)
    r = lib().orc_project_strain_f64(_ptr(F), float(mu), float(lam), int(plasticity), float(friction_angle),
                                     float(yield_stress))
    return F.reshape(3, 3), bool(r)


def project_cam_clay(F, mu, lam, friction_angle=35.0, beta=2.0, xi=1.0, logJp=-0.04):
    """Non-associative Cam-clay return mapping of one F (FP64). Returns (F', logJp', changed)."""
    F = np.array(F, dtype=np.float64).reshape(9)
    lj = ctypes.c_double(float(logJp))
    r = lib().orc_project_cam_clay_f64(_ptr(F), float(mu), float(lam), float(friction_angle), float(beta),
                                       float(xi), ctypes.byref(lj))
    return F.reshape(3, 3), lj.value, bool(r)


class Collider(ctypes.Structure):
    """zmpm_collider's C layout with the SDF samples on the host."""
    _fields_ = [("type", ctypes.c_int32), ("shape", ctypes.c_int32), ("scene", ctypes.c_int32),
                ("dims", ctypes.c_int32 * 3), ("shape_params", ctypes.c_float * 6), ("sdf", ctypes.c_void_p),
                ("origin", ctypes.c_float * 3), ("spacing", ctypes.c_float), ("translation", ctypes.c_float * 3),
                ("translation_rate", ctypes.c_float * 3), ("scale", ctypes.c_float), ("scale_rate", ctypes.c_float),
                ("ypr_degrees", ctypes.c_float * 3), ("angular_velocity", ctypes.c_float * 3),
                ("leaf_origins", ctypes.c_void_p), ("nleaves", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def make_collider(d, keep):
    """d: dict with the zmpm_collider fields (sdf: host float32 array (nx, ny, nz); for the
    SDF leaves shape sdf (nleaves, 8, 8, 8) and leaf_origins (nleaves, 3) int32)."""
    c = Collider()
    c.type, c.shape, c.scene = int(d.get("type", 0)), int(d.get("shape", 0)), int(d.get("scene", 0))
    p = list(d.get("params", ())) + [0.0] * 6
    for k in range(6):
        c.shape_params[k] = float(p[k])
    if d.get("sdf") is not None:
        a = np.ascontiguousarray(d["sdf"], dtype=np.float32)
        keep.append(a)
        c.sdf = ctypes.c_void_p(a.ctypes.data)
        for k in range(3):
            c.dims[k] = a.shape[k] if a.ndim == 3 else 0
    if d.get("leaf_origins") is not None:
        o = np.ascontiguousarray(d["leaf_origins"], dtype=np.int32).reshape(-1, 3)
        keep.append(o)
        c.leaf_origins = ctypes.c_void_p(o.ctypes.data)
        c.nleaves = o.shape[0]
    for name in ("origin", "translation", "translation_rate", "ypr_degrees", "angular_velocity"):
        vals = d.get(name, (0.0, 0.0, 0.0))
        arr = getattr(c, name)
        for k in range(3):
            arr[k] = float(vals[k])
    c.spacing = float(d.get("spacing", 1.0))
    c.scale = float(d.get("scale", 1.0))
    c.scale_rate = float(d.get("scale_rate", 0.0))
    return c


def shape_phi(d, X):
    """The collider's shape alone (FP64): phi and its gradient at X in the collider frame."""
    keep = []
    c = make_collider(d, keep)
    X = np.ascontiguousarray(X, dtype=np.float64)
    g = np.zeros(3)
    phi = lib().orc_shape_phi_f64(ctypes.byref(c), _ptr(X), _ptr(g))
    return phi, g


def resolve_collision(d, x, v):
    """One collider response (FP64) at position x on velocity v. Returns (v', inside)."""
    keep = []
    c = make_collider(d, keep)
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    r = lib().orc_resolve_collision_f64(ctypes.byref(c), _ptr(x), _ptr(v))
    return v, bool(r)


def uniform_coord(x, bv_min, bv_max):
    """AABBBox::getUniformCoord as orderByMortonCode uses it (projects/CUDA/Structures.hpp:425):
    (x - min) / (max - min) per axis, FP32 with a correctly rounded division."""
    x = np.asarray(x, np.float32)
    lo = np.asarray(bv_min, np.float32)[:, None]
    ext = np.asarray(bv_max, np.float32)[:, None] - lo
    return ((x - lo) / ext).astype(np.float32)


def morton_code3(coord):
    """The u32 3D Morton code of orderByMortonCode (projects/CUDA/Structures.hpp:426):
    per axis q = trunc(min(max(c * 1024, 0), 1023)) (10 bits), interleaved so that bit b of
    x, y, z lands at 3b + 2, 3b + 1, 3b (Karras 2012, "Maximizing parallelism in the
    construction of BVHs, octrees, and k-d trees", section 4).  Restated here bit by bit;
    zpc's morton_code<3> is absent, so parity against zpc itself is unpinned."""
    c = np.asarray(coord, np.float32)
    q = np.minimum(np.maximum(c * np.float32(1024.0), np.float32(0.0)), np.float32(1023.0)).astype(np.uint32)
    code = np.zeros(c.shape[1], np.uint32)
    for b in range(10):
        for axis in range(3):
            code |= ((q[axis] >> np.uint32(b)) & np.uint32(1)) << np.uint32(3 * b + 2 - axis)
    return code


def order_by_morton_code(x, bv_min, bv_max):
    """ZenoParticles::orderByMortonCode (projects/CUDA/Structures.hpp:406-435) on positions x (3, n):
    the codes, orderedToOriginal (a stable sort of (code, index): radix_sort_pair,
    Structures.hpp:430-431) and its inverse originalToOrdered (Structures.hpp:433-435).  The
    caller permutes the channels with orderedToOriginal (Structures.hpp:438-443: ordered
    particle k takes every channel of original particle orderedToOriginal[k])."""
    codes = morton_code3(uniform_coord(x, bv_min, bv_max))
    o2o = np.argsort(codes, kind="stable").astype(np.int32)
    inv = np.empty_like(o2o)
    inv[o2o] = np.arange(o2o.size, dtype=np.int32)
    return codes, o2o, inv
