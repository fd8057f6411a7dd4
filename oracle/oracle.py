"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module.  The product (lenia_amd) never does.  The oracle restates the reference step
(lenia.cpp:435-1046, reintegration.cpp:13-1410, transposition.cpp:6-995) in C; see
oracle/lenia_oracle.h for what pins each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblenia_oracle.so")

MAX_K, MAX_C, MAX_BETA = 16, 4, 12

MODE = {"srt": 0, "rt": 1, "adv": 2}
BORDER = {"torus": 0, "wall": 1, "none": 2}
GRADIENT = {"central": 0, "sobel": 1}
GROWTH = {"none": 0, "rt": 1, "transposition": 2}

Q_SOBEL_PERMUTE, Q_PREALPHA_INDEX, Q_SUM_RESET, Q_FACE_HALO = 1, 2, 4, 8
Q_TRANSPORT_TORUS, Q_RT_3CH, Q_GROWTH_LUT, Q_RT_SNORM = 16, 32, 64, 128
Q_CLAMP01, Q_CD_COPY, Q_SOFTCLIP_TR = 256, 512, 1024
Q_ALL = 0x7FF

F_A, F_GA, F_S, F_V, F_H, F_B, F_MU, F_USUM, F_F, F_NA, F_NU, F_TT, F_U = range(13)


class Params(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("border", C.c_int), ("sobel_border", C.c_int),
        ("gradient", C.c_int), ("growth_type", C.c_int),
        ("mu_vel", C.c_int), ("adv_vel", C.c_int), ("blur", C.c_int),
        ("dt", C.c_float), ("sigma", C.c_float), ("dd", C.c_float), ("s_norm", C.c_float),
        ("theta_a", C.c_float), ("gta", C.c_float), ("sta", C.c_float), ("v_damp", C.c_float),
        ("asymptotic", C.c_int * MAX_C), ("soft_clip", C.c_int * MAX_C),
        ("osc_la", C.c_int), ("osc_flow", C.c_int), ("osc_div_flow", C.c_int), ("osc_vel", C.c_int),
        ("osc_la_amt", C.c_float), ("osc_flow_amt", C.c_float), ("osc_div_flow_amt", C.c_float),
        ("quirks", C.c_uint32),
    ]


class Kernel(C.Structure):
    _fields_ = [
        ("c0", C.c_int), ("c1", C.c_int), ("radius", C.c_int), ("beta_size", C.c_int),
        ("b", C.c_float * MAX_BETA), ("r_a", C.c_float), ("m", C.c_float), ("s", C.c_float),
        ("h", C.c_float),
    ]


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_native() -> str:
    """Builds the host-tuned copy (-O3 -march=native, same results) on this machine and
    makes lib() load it; call before the first lib() of the process."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("oracle library already loaded")
    subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
    _LIB_PATH = os.path.join(_HERE, "build", "native", "liblenia_oracle.so")
    return _LIB_PATH


def set_threads(n: int) -> int:
    """Host threads of the checker's step loops (results do not depend on it)."""
    return lib().lo_set_threads(int(n))


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        for pfx, T in (("lo_f32_", C.c_float), ("lo_f64_", C.c_double)):
            P = C.POINTER(T)
            f = getattr(L, pfx + "create")
            f.restype = C.c_void_p
            f.argtypes = [C.c_int] * 5 + [C.POINTER(Params), C.POINTER(Kernel)]
            getattr(L, pfx + "destroy").argtypes = [C.c_void_p]
            for n in ("set_field", "get_field"):
                g = getattr(L, pfx + n)
                g.argtypes = [C.c_void_p, C.c_int, P]
                g.restype = C.c_int
            getattr(L, pfx + "step").argtypes = [C.c_void_p, C.c_int]
            getattr(L, pfx + "step_phases").argtypes = [C.c_void_p, C.c_int]
            m = getattr(L, pfx + "mass")
            m.argtypes = [C.c_void_p]
            m.restype = C.c_double
            getattr(L, pfx + "fft1").argtypes = [C.c_int, P, C.c_int]
            getattr(L, pfx + "rfft3").argtypes = [C.c_int] * 3 + [P, P]
            getattr(L, pfx + "irfft3").argtypes = [C.c_int] * 3 + [P, P]
            getattr(L, pfx + "kernel_field").argtypes = [C.c_int] * 3 + [C.POINTER(Kernel), P]
            getattr(L, pfx + "kernel_spectrum").argtypes = [C.c_int] * 3 + [P, P]
            getattr(L, pfx + "blur_field").argtypes = [C.c_int] * 3 + [P]
            getattr(L, pfx + "bell_table").argtypes = [C.c_float, C.c_float, P]
            getattr(L, pfx + "mc_case").argtypes = [C.c_int] * 3 + [P, T, C.c_int, C.POINTER(C.c_uint8)]
        L.lo_set_threads.argtypes = [C.c_int]
        L.lo_set_threads.restype = C.c_int
        L.lo_edge_mask.argtypes = [C.c_int]
        L.lo_edge_mask.restype = C.c_int
        L.lo_canon_exp.restype = C.c_double
        L.lo_canon_exp.argtypes = [C.c_double]
        FP = C.POINTER(C.c_float)
        L.lo_mesh.argtypes = ([C.c_int] * 5 + [FP] * 3 + [C.c_float, C.c_float, C.c_int, C.POINTER(C.c_int),
                              C.POINTER(C.c_uint32), C.c_void_p, C.c_int64, C.POINTER(C.c_uint32), C.c_int64,
                              C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
        L.lo_mesh.restype = C.c_int
        _lib = L
    return _lib


def _dt(prec):
    return (np.float32, "lo_f32_", C.c_float) if prec == 32 else (np.float64, "lo_f64_", C.c_double)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def make_params(cfg: dict) -> Params:
    p = Params()
    p.mode = MODE[cfg.get("mode", "rt")]
    p.border = BORDER[cfg.get("border", "torus")]
    p.sobel_border = BORDER[cfg.get("sobel_border", "torus")]
    p.gradient = GRADIENT[cfg.get("gradient", "sobel")]
    p.growth_type = GROWTH[cfg.get("growth", "rt")]
    p.mu_vel = int(cfg.get("mu_vel", False))
    p.adv_vel = int(cfg.get("adv_vel", True))
    p.blur = int(cfg.get("blur", False))
    for k, d in (("dt", 0.2), ("sigma", 0.65), ("dd", 1.0), ("s_norm", 0.0), ("theta_a", 2.0),
                 ("gta", 0.1), ("sta", 0.1), ("v_damp", 0.9), ("osc_la_amt", -0.001),
                 ("osc_flow_amt", -0.001), ("osc_div_flow_amt", -0.001)):
        setattr(p, k, float(cfg.get(k, d)))
    for c in range(MAX_C):
        p.asymptotic[c] = int(cfg.get("asymptotic", [0] * MAX_C)[c])
        p.soft_clip[c] = int(cfg.get("soft_clip", [0] * MAX_C)[c])
    for k in ("osc_la", "osc_flow", "osc_div_flow", "osc_vel"):
        setattr(p, k, int(cfg.get(k, False)))
    p.quirks = int(cfg.get("quirks", Q_ALL))
    return p


def make_kernels(kernels: list) -> C.Array:
    arr = (Kernel * max(1, len(kernels)))()
    for i, k in enumerate(kernels):
        arr[i].c0, arr[i].c1 = int(k["c0"]), int(k["c1"])
        arr[i].radius, arr[i].beta_size = int(k["radius"]), int(k["beta_size"])
        for j, b in enumerate(k["b"]):
            arr[i].b[j] = float(b)
        arr[i].r_a, arr[i].m, arr[i].s, arr[i].h = (float(k[n]) for n in ("r_a", "m", "s", "h"))
    return arr


class World:
    """One oracle world (float or double)."""

    def __init__(self, shape, C_, kernels, cfg, prec=32):
        self.W, self.H, self.D = shape
        self.C, self.K = C_, len(kernels)
        self.N = self.W * self.H * self.D
        self.dtype, self.pfx, self.ctype = _dt(prec)
        self._p = make_params(cfg)
        self._k = make_kernels(kernels)
        L = lib()
        self._h = getattr(L, self.pfx + "create")(self.W, self.H, self.D, self.C, self.K,
                                                   C.byref(self._p), self._k)
        if not self._h:
            raise ValueError("oracle rejected world shape/params")

    def __del__(self):
        if getattr(self, "_h", None):
            getattr(lib(), self.pfx + "destroy")(self._h)
            self._h = None

    def _size(self, field):
        C_, N = self.C, self.N
        return {F_A: C_ * N, F_GA: C_ * N, F_S: N, F_V: 3 * C_ * N, F_H: C_ * N, F_B: C_ * N,
                F_MU: 3 * C_ * N, F_USUM: C_ * N, F_F: 3 * C_ * N, F_NA: 3 * N, F_NU: 3 * C_ * N,
                F_TT: C_ * N, F_U: max(1, self.K) * N}[field]

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).reshape(-1)
        assert a.size == self._size(field)
        rc = getattr(lib(), self.pfx + "set_field")(self._h, field, _ptr(a, self.ctype))
        assert rc == 0

    def get(self, field):
        a = np.empty(self._size(field), dtype=self.dtype)
        rc = getattr(lib(), self.pfx + "get_field")(self._h, field, _ptr(a, self.ctype))
        assert rc == 0
        return a

    def step(self, n=1):
        getattr(lib(), self.pfx + "step")(self._h, int(n))

    def step_phases(self, n):
        """Run the first n phases of one step (7: up to and including reintegration)."""
        getattr(lib(), self.pfx + "step_phases")(self._h, int(n))

    def mass(self):
        return getattr(lib(), self.pfx + "mass")(self._h)


def fft1(x: np.ndarray, inverse=False, prec=32):
    dtype, pfx, ct = _dt(prec)
    a = np.empty(2 * x.size, dtype=dtype)
    a[0::2], a[1::2] = x.real, x.imag
    getattr(lib(), pfx + "fft1")(x.size, _ptr(a, ct), int(inverse))
    return a[0::2] + 1j * a[1::2]


def rfft3(f: np.ndarray, prec=32):
    """f: [D][H][W] real -> [D][H][W/2+1] complex."""
    D, H, W = f.shape
    dtype, pfx, ct = _dt(prec)
    a = np.ascontiguousarray(f, dtype=dtype)
    out = np.empty(2 * D * H * (W // 2 + 1), dtype=dtype)
    getattr(lib(), pfx + "rfft3")(W, H, D, _ptr(a, ct), _ptr(out, ct))
    return (out[0::2] + 1j * out[1::2]).reshape(D, H, W // 2 + 1)


def irfft3(F: np.ndarray, W: int, prec=32):
    D, H, Wh = F.shape
    dtype, pfx, ct = _dt(prec)
    a = np.empty(2 * F.size, dtype=dtype)
    a[0::2], a[1::2] = F.real.reshape(-1), F.imag.reshape(-1)
    out = np.empty(D * H * W, dtype=dtype)
    getattr(lib(), pfx + "irfft3")(W, H, D, _ptr(a, ct), _ptr(out, ct))
    return out.reshape(D, H, W)


def kernel_field(shape, kernel: dict, prec=32):
    W, H, D = shape
    dtype, pfx, ct = _dt(prec)
    out = np.empty(W * H * D, dtype=dtype)
    k = make_kernels([kernel])
    getattr(lib(), pfx + "kernel_field")(W, H, D, k, _ptr(out, ct))
    return out.reshape(D, H, W)


def kernel_spectrum(field: np.ndarray, prec=32):
    D, H, W = field.shape
    dtype, pfx, ct = _dt(prec)
    a = np.ascontiguousarray(field, dtype=dtype)
    out = np.empty((D // 2 + 1) * (H // 2 + 1) * (W // 2 + 1), dtype=dtype)
    getattr(lib(), pfx + "kernel_spectrum")(W, H, D, _ptr(a, ct), _ptr(out, ct))
    return out.reshape(D // 2 + 1, H // 2 + 1, W // 2 + 1)


def blur_field(shape, prec=32):
    W, H, D = shape
    dtype, pfx, ct = _dt(prec)
    out = np.empty(W * H * D, dtype=dtype)
    getattr(lib(), pfx + "blur_field")(W, H, D, _ptr(out, ct))
    return out.reshape(D, H, W)


def bell_table(m, s, prec=32):
    dtype, pfx, ct = _dt(prec)
    out = np.empty(256, dtype=dtype)
    getattr(lib(), pfx + "bell_table")(float(m), float(s), _ptr(out, ct))
    return out


def mc_case(field: np.ndarray, iso=0.5, face_only=True, prec=32):
    D, H, W = field.shape
    dtype, pfx, ct = _dt(prec)
    a = np.ascontiguousarray(field, dtype=dtype)
    out = np.empty(W * H * D, dtype=np.uint8)
    getattr(lib(), pfx + "mc_case")(W, H, D, _ptr(a, ct), ct(iso), int(face_only),
                                    out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.reshape(D, H, W)


VERTEX_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("nx", "<f4"), ("ny", "<f4"),
                         ("nz", "<f4"), ("color", "<u4")])


def mesh(fields, tri_table, iso=0.5, cell=3.5, face_only=True, colors=None):
    """Marching-cubes mesh of 1-3 fields (A, U_SUM, H), each [C][D][H][W] float32, after
    generate_mesh_v1/v2/v3 (mesh.cpp:34-2036) with one section: returns (vertices as
    VERTEX_DTYPE records, uint32 indices)."""
    F = len(fields)
    assert 1 <= F <= 3
    arrs = [np.ascontiguousarray(f, dtype=np.float32) for f in fields]
    Cn, D, H, W = arrs[0].shape
    tri = np.ascontiguousarray(tri_table, dtype=np.int32).reshape(256, 16)
    col = np.ascontiguousarray(colors if colors is not None else [0xFFFFFFFF] * Cn, dtype=np.uint32)
    FP = C.POINTER(C.c_float)
    ptrs = [a.ctypes.data_as(FP) for a in arrs] + [None] * (3 - F)
    nv, ni = C.c_int64(), C.c_int64()
    L = lib()
    r = L.lo_mesh(W, H, D, Cn, F, *ptrs, iso, cell, int(face_only), tri.ctypes.data_as(C.POINTER(C.c_int)),
                  col.ctypes.data_as(C.POINTER(C.c_uint32)), None, 0, None, 0, C.byref(nv), C.byref(ni))
    assert r in (0, -1), r
    verts = np.zeros(nv.value, dtype=VERTEX_DTYPE)
    idx = np.zeros(ni.value, dtype=np.uint32)
    r = L.lo_mesh(W, H, D, Cn, F, *ptrs, iso, cell, int(face_only), tri.ctypes.data_as(C.POINTER(C.c_int)),
                  col.ctypes.data_as(C.POINTER(C.c_uint32)), verts.ctypes.data_as(C.c_void_p), nv.value,
                  idx.ctypes.data_as(C.POINTER(C.c_uint32)), ni.value, C.byref(nv), C.byref(ni))
    assert r == 0, r
    return verts, idx


def canon_exp(x) -> np.ndarray:
    """The checker's exp for the analytic growth bell, elementwise over x (float64)."""
    f = lib().lo_canon_exp
    return np.array([f(float(v)) for v in np.ravel(x)], dtype=np.float64).reshape(np.shape(x))


def edge_mask(case_index: int) -> int:
    return lib().lo_edge_mask(int(case_index))
