"""Python front-end of the MI355X Flow Lenia engine (over the C-ABI in lenia_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class Engine:
    """A batch of worlds resident in HBM, advanced together on one HIP stream.

    >>> eng = Engine(); wid = eng.add_world(shape, C, kernels, cfg); eng.commit()
    >>> eng.set(wid, "A", A0); eng.step(100); A = eng.get(wid, "A")
    """

    FIELDS = {"A": L.F_A, "GA": L.F_GA, "S": L.F_S, "V": L.F_V, "H": L.F_H, "B": L.F_B,
              "MU": L.F_MU, "USUM": L.F_USUM}

    def __init__(self, device: int = 0, keep_velocity=False, keep_usum=False, graph=True,
                 profile=False, autotune=True, schedule=None, keep_mu=False):
        self._L = L.load()
        h = C.c_void_p()
        L.check(self._L.lh_create(device, C.byref(h)), "lh_create")
        self._h = h
        self.worlds = []
        L.check(self._L.lh_set_option(self._h, L.OPT_KEEP_VELOCITY, int(keep_velocity)), "keep_velocity")
        L.check(self._L.lh_set_option(self._h, L.OPT_KEEP_USUM, int(keep_usum)), "keep_usum")
        L.check(self._L.lh_set_option(self._h, L.OPT_KEEP_MU, int(keep_mu)), "keep_mu")
        self.set_option(L.OPT_USE_GRAPH, graph)
        self.set_option(L.OPT_PROFILE, profile)
        self.set_option(L.OPT_AUTOTUNE, autotune)
        if schedule is not None:
            # forced sweep segments (local, transport), tiled transport, growth (0 default,
            # 1 spec_z + inv_growth, 2 fused)
            zl, zt, tiled = schedule[:3]
            growth = schedule[3] if len(schedule) > 3 else 0
            self.set_option(L.OPT_SCHEDULE, int(zl) | int(zt) << 8 | int(bool(tiled)) << 16 | int(growth) << 17)

    def set_option(self, opt, value):
        L.check(self._L.lh_set_option(self._h, opt, int(value)), "lh_set_option")

    def close(self):
        if getattr(self, "_h", None):
            self._L.lh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_world(self, shape, C_, kernels, cfg, conv="fft") -> int:
        """conv: growth convolution, "fft" (fp32, reference-exact) or a matrix-core
        DFT-as-GEMM variant "bf16x3", "fp16x3" (split hi+lo), "bf16", "fp16"."""
        W, H, D = shape
        s = L.Shape(W, H, D, C_, len(kernels))
        p = L.make_params(cfg)
        k = L.make_kernels(kernels)
        wid = C.c_int()
        L.check(self._L.lh_add_world(self._h, C.byref(s), C.byref(p), k, C.byref(wid)), "lh_add_world")
        if conv != "fft":
            L.check(self._L.lh_set_world_conv(self._h, wid.value, L.CONV[conv]), "lh_set_world_conv")
        self.worlds.append((tuple(shape), C_, len(kernels)))
        return wid.value

    def add_world_from_checkpoint(self, image: bytes) -> int:
        """Adds a world with the shape, parameters, kernels and convolution recorded in a
        checkpoint image (its state is loaded with restore() after commit)."""
        ck = read_checkpoint(image)
        wid = C.c_int()
        L.check(self._L.lh_add_world(self._h, C.byref(ck["shape"]), C.byref(ck["params"]), ck["kernels"],
                                     C.byref(wid)), "lh_add_world")
        if ck["conv"] != L.CONV["fft"]:
            L.check(self._L.lh_set_world_conv(self._h, wid.value, ck["conv"]), "lh_set_world_conv")
        s = ck["shape"]
        self.worlds.append(((s.W, s.H, s.D), s.C, s.K))
        return wid.value

    def commit(self):
        L.check(self._L.lh_commit(self._h), "lh_commit")

    def checkpoint(self, wid) -> np.ndarray:
        """The world's checkpoint image (lh_checkpoint_save: header, parameters, kernels,
        A, GA, S, V) as a uint8 array (bytes-like: write it to a file as is)."""
        n = C.c_size_t()
        L.check(self._L.lh_checkpoint_size(self._h, wid, C.byref(n)), "lh_checkpoint_size")
        buf = np.empty(n.value, dtype=np.uint8)
        L.check(self._L.lh_checkpoint_save(self._h, wid, buf.ctypes.data, n.value), "lh_checkpoint_save")
        return buf

    def restore(self, wid, image):
        """Loads a checkpoint image (any bytes-like object) into a world with the same
        shape, parameters, kernels and convolution (lh_checkpoint_load)."""
        b = np.ascontiguousarray(np.frombuffer(image, dtype=np.uint8))
        L.check(self._L.lh_checkpoint_load(self._h, wid, b.ctypes.data, b.size), "lh_checkpoint_load")

    def seed_random(self, first_world, n_worlds, seed, first_set=0):
        """Seeds worlds first_world.. with the Gaussian blobs of sampled sets first_set..
        (lh_seed_random: A from five blobs, GA, S, V and step count cleared)."""
        L.check(self._L.lh_seed_random(self._h, int(first_world), int(n_worlds), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                       int(first_set)), "lh_seed_random")

    SNAP = {"A": 1, "H": 2, "USUM": 4}

    def snapshot(self, slot, first_world=0, n_worlds=None, fields=("A",)):
        """Enqueues a snapshot of fields of worlds first_world.. into slot 0 or 1, ordered
        after the steps enqueued so far (lh_snapshot); returns at once."""
        n = len(self.worlds) - first_world if n_worlds is None else n_worlds
        mask = sum(self.SNAP[f] for f in set(fields))
        L.check(self._L.lh_snapshot(self._h, int(slot), int(first_world), int(n), mask), "lh_snapshot")

    def snapshot_ready(self, slot) -> bool:
        r = C.c_int()
        L.check(self._L.lh_snapshot_ready(self._h, int(slot), C.byref(r)), "lh_snapshot_ready")
        return bool(r.value)

    def snapshot_read(self, slot, wid, field="A") -> np.ndarray:
        """One field of one world of a snapshot; waits for that snapshot only."""
        a = np.empty(self._count(wid, field), dtype=np.float32)
        L.check(self._L.lh_snapshot_read(self._h, int(slot), int(wid), self.FIELDS[field],
                                         a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "lh_snapshot_read")
        return a

    def steps(self, wid) -> int:
        n = C.c_uint64()
        L.check(self._L.lh_world_steps(self._h, wid, C.byref(n)), "lh_world_steps")
        return n.value

    def _count(self, wid, field):
        (W, H, D), C_, _ = self.worlds[wid]
        N = W * H * D
        return {"A": C_ * N, "GA": C_ * N, "S": N, "V": 3 * C_ * N, "H": C_ * N, "B": C_ * N,
                "MU": 3 * C_ * N, "USUM": C_ * N}[field]

    def set(self, wid, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1)
        if a.size != self._count(wid, field):
            raise ValueError(f"{field}: expected {self._count(wid, field)} floats, got {a.size}")
        L.check(self._L.lh_set_field(self._h, wid, self.FIELDS[field],
                                     a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "lh_set_field")

    def get(self, wid, field) -> np.ndarray:
        a = np.empty(self._count(wid, field), dtype=np.float32)
        L.check(self._L.lh_get_field(self._h, wid, self.FIELDS[field],
                                     a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "lh_get_field")
        return a

    def device_ptr(self, wid, field) -> int:
        p = C.POINTER(C.c_float)()
        n = C.c_size_t()
        L.check(self._L.lh_field_device_ptr(self._h, wid, self.FIELDS[field], C.byref(p), C.byref(n)),
                "lh_field_device_ptr")
        return C.cast(p, C.c_void_p).value

    def state_updated(self):
        L.check(self._L.lh_state_updated(self._h), "lh_state_updated")

    def step(self, n=1):
        """Enqueues n steps (up to 256 steps are one captured-graph launch)."""
        L.check(self._L.lh_step(self._h, int(n)), "lh_step")

    def prepare_steps(self, n):
        """Captures the graphs step(n) launches, so the first such call starts at once."""
        L.check(self._L.lh_prepare_steps(self._h, int(n)), "lh_prepare_steps")

    def synchronize(self):
        L.check(self._L.lh_synchronize(self._h), "lh_synchronize")

    def stream(self) -> int:
        return self._L.lh_stream(self._h)

    def mass(self) -> np.ndarray:
        out = np.empty(len(self.worlds), dtype=np.float64)
        L.check(self._L.lh_world_mass(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), out.size),
                "lh_world_mass")
        return out

    def moments(self) -> np.ndarray:
        """[n_worlds, 7] FP64: mass, then (sum A cos, sum A sin) for x, y, z."""
        out = np.empty((len(self.worlds), 7), dtype=np.float64)
        L.check(self._L.lh_world_moments(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), len(self.worlds)),
                "lh_world_moments")
        return out

    def centre_of_mass(self) -> np.ndarray:
        """[n_worlds, 3] torus centre of mass (x, y, z) in cells from moments()."""
        m = self.moments()
        ext = np.array([s[0] for s, _, _ in self.worlds]), np.array([s[1] for s, _, _ in self.worlds]), \
            np.array([s[2] for s, _, _ in self.worlds])
        out = np.empty((len(self.worlds), 3))
        for a in range(3):
            ang = np.arctan2(m[:, 2 + 2 * a], m[:, 1 + 2 * a])
            out[:, a] = np.mod(ang / (2 * np.pi) * ext[a] - 0.5, ext[a])
        return out

    def mc_case(self, wid, field="A", channel=0, iso=0.5, face_only=True, noalloc=False) -> np.ndarray:
        """Case index bytes [D][H][W]; noalloc=True calls lh_mc_case_noalloc (static scratch)."""
        (W, H, D), _, _ = self.worlds[wid]
        out = np.empty(W * H * D, dtype=np.uint8)
        name = "lh_mc_case_noalloc" if noalloc else "lh_mc_case"
        L.check(getattr(self._L, name)(self._h, wid, self.FIELDS[field], channel, float(iso), int(face_only),
                                       out.ctypes.data_as(C.POINTER(C.c_uint8))), name)
        return out.reshape(D, H, W)

    def mesh(self, wid, view=1, iso=0.5, cell_size=3.5, face_only=True, tri_table=None, colors=None,
             snapshot_slot=None):
        """Marching-cubes mesh of world wid for view 1 (A), 2 (A, U_SUM) or 3 (A, U_SUM, H), as
        the renderer's mesh threads build it (mesh.cpp:34, 434, 1115): returns (vertices as
        VERTEX_DTYPE records, uint32 indices).  tri_table: 256x16 ints, default mc_table().
        snapshot_slot: mesh the fields of that snapshot (lh_mesh_snapshot) instead of the
        current state, while later steps keep running."""
        tri = None
        if tri_table is not None:
            tri = np.ascontiguousarray(tri_table, dtype=np.int32).reshape(256, 16)
        col = None
        if colors is not None:
            col = np.ascontiguousarray(colors, dtype=np.uint32)
            if col.size < self.worlds[wid][1]:
                raise ValueError("one colour per channel")
        tp = tri.ctypes.data_as(C.POINTER(C.c_int)) if tri is not None else None
        cp = col.ctypes.data_as(C.POINTER(C.c_uint32)) if col is not None else None
        nv, ni = C.c_int64(), C.c_int64()
        args = (self._h, wid, int(view), float(iso), float(cell_size), int(face_only), tp, cp)
        fn, what = self._L.lh_mesh, "lh_mesh"
        if snapshot_slot is not None:
            args = (self._h, int(snapshot_slot)) + args[1:]
            fn, what = self._L.lh_mesh_snapshot, "lh_mesh_snapshot"
        # one pass when the buffers of the last mesh of this size class are large enough,
        # else a second pass into buffers of the counts the first one reported
        key = (wid, int(view))
        cap_v, cap_i = getattr(self, "_mesh_caps", {}).get(key, (0, 0))
        for _ in range(2):
            verts = np.empty(cap_v, dtype=VERTEX_DTYPE)
            idx = np.empty(cap_i, dtype=np.uint32)
            rc = fn(*args, verts.ctypes.data_as(C.c_void_p) if cap_v else None, cap_v,
                                 idx.ctypes.data_as(C.POINTER(C.c_uint32)) if cap_i else None, cap_i,
                                 C.byref(nv), C.byref(ni))
            fits = nv.value <= cap_v and ni.value <= cap_i
            if rc == L.OK and fits and (cap_v or not nv.value):
                break
            if fits:
                L.check(rc, what)
            cap_v, cap_i = max(cap_v, int(nv.value * 1.25) + 64), max(cap_i, int(ni.value * 1.25) + 192)
        if not hasattr(self, "_mesh_caps"):
            self._mesh_caps = {}
        self._mesh_caps[key] = (cap_v, cap_i)
        return verts[:nv.value], idx[:ni.value]

    def profile(self) -> dict:
        out = {}
        for i in range(self._L.lh_profile_count(self._h)):
            name, n, ms = C.c_char_p(), C.c_long(), C.c_double()
            L.check(self._L.lh_profile_entry(self._h, i, C.byref(name), C.byref(n), C.byref(ms)),
                    "lh_profile_entry")
            out[name.value.decode()] = (n.value, ms.value)
        return out

    def calibration_copy(self, nbytes: int):
        L.check(self._L.lh_calibration_copy(self._h, int(nbytes)), "lh_calibration_copy")

    def profile_reset(self):
        L.check(self._L.lh_profile_reset(self._h), "lh_profile_reset")


def shape_supported(W, H, D) -> int:
    """1: built into the library; 2: available as a shape module (build one with
    `python -m lenia_amd.build --shape WxHxD`); 0: neither."""
    return int(L.load().lh_shape_supported(W, H, D))


VERTEX_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("nx", "<f4"), ("ny", "<f4"),
                         ("nz", "<f4"), ("color", "<u4")])  # lh_vertex, the reference's 28-byte Vertex


def mc_table() -> np.ndarray:
    """The library's default marching-cubes triangulation table, int32 [256][16] with -1
    terminated rows of edge triples (host code, no GPU needed)."""
    out = np.empty((256, 16), dtype=np.int32)
    L.check(L.load().lh_mc_table(out.ctypes.data_as(C.POINTER(C.c_int))), "lh_mc_table")
    return out


def sample_random_raw(seed, first, n, channels, device=0):
    """lh_sample_random into fresh host buffers: (lh_kernel array [n * LH_MAX_K], kernel
    counts [n])."""
    lib = L.load()
    ks = (L.Kernel * max(1, n * L.MAX_K))()
    cnt = np.zeros(max(1, n), dtype=np.int32)
    L.check(lib.lh_sample_random(int(device), int(seed) & 0xFFFFFFFFFFFFFFFF, int(first), int(n), int(channels),
                                 ks, cnt.ctypes.data_as(C.POINTER(C.c_int))), "lh_sample_random")
    return ks, cnt[:n]


def sample_random(seed, first, n, channels, device=0) -> list:
    """Parameter sets first .. first+n-1 sampled on the GPU with generate_random's ranges
    (lh_sample_random): one list of kernel dicts (add_world's format) per set."""
    ks, cnt = sample_random_raw(seed, first, n, channels, device)
    out = []
    for j in range(n):
        sets = []
        for i in range(cnt[j]):
            k = ks[j * L.MAX_K + i]
            sets.append(dict(c0=k.c0, c1=k.c1, radius=k.radius, beta_size=k.beta_size,
                             b=[k.b[e] for e in range(k.beta_size)], r_a=k.r_a, m=k.m, s=k.s, h=k.h))
        out.append(sets)
    return out


def read_checkpoint(image: bytes) -> dict:
    """Parses a checkpoint image (layout of lh_checkpoint_header in lenia_hip.h) without
    the engine: header fields, the lh_params / lh_kernel structures and the state fields
    as float32 arrays (A [C][N], GA [C][N], S [N], V [C][3][N])."""
    b = bytes(image)  # one copy; the arrays below are independent of the caller's buffer
    hs = C.sizeof(L.CheckpointHeader)
    if len(b) < hs:
        raise ValueError("checkpoint image truncated")
    h = L.CheckpointHeader.from_buffer_copy(b[:hs])
    if h.magic != b"LHCK" or h.version != L.CHECKPOINT_VERSION or h.header_bytes != hs:
        raise ValueError("not a version-1 checkpoint image")
    if h.params_bytes != C.sizeof(L.Params) or h.kernel_bytes != C.sizeof(L.Kernel):
        raise ValueError("checkpoint written with a different structure layout")
    s = h.shape
    N = s.W * s.H * s.D
    sizes = {"A": s.C * N, "GA": s.C * N, "S": N, "V": 3 * s.C * N}
    off = hs + h.params_bytes + s.K * h.kernel_bytes
    if h.payload_bytes != 4 * sum(sizes.values()) or len(b) != off + h.payload_bytes:
        raise ValueError("checkpoint size does not match its header")
    params = L.Params.from_buffer_copy(b[hs:hs + h.params_bytes])
    kernels = (L.Kernel * s.K).from_buffer_copy(b[hs + h.params_bytes:off])
    fields = {}
    for name, n in sizes.items():
        fields[name] = np.frombuffer(b, dtype="<f4", count=n, offset=off).copy()
        off += 4 * n
    shape = L.Shape(s.W, s.H, s.D, s.C, s.K)
    return dict(version=h.version, steps=h.steps, conv=h.conv, shape=shape, params=params, kernels=kernels,
                fields=fields)
