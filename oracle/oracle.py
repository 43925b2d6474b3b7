"""ctypes binding of the CPU restatement in pm_oracle.c.

TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation and the
timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpm_oracle.so")

MAX_SRC = 8
DOMAIN_INIT = 0x494E4954
DOMAIN_REFINE = 0x52464E45
FLAG_NP, FLAG_WZ, FLAG_EDGE, FLAG_DEN = 1, 2, 4, 8
COST_FP32, COST_FP64 = 0, 1

_u8p = ctypes.POINTER(ctypes.c_uint8)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class OrcView(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int), ("H", ctypes.c_int), ("S", ctypes.c_int),
        ("ref", _u8p), ("src", _u8p * MAX_SRC),
        ("Kinv", ctypes.c_double * 9),
        ("A", (ctypes.c_double * 9) * MAX_SRC),
        ("b", (ctypes.c_double * 3) * MAX_SRC),
        ("dmin", ctypes.c_float), ("dmax", ctypes.c_float),
        ("radius", ctypes.c_int), ("stride", ctypes.c_int), ("top_k", ctypes.c_int),
        ("ref_gray", _f32p), ("src_gray", _f32p * MAX_SRC),
    ]


class OrcGeomSrc(ctypes.Structure):
    _fields_ = [("K", ctypes.c_float * 9), ("R", ctypes.c_float * 9), ("t", ctypes.c_float * 3),
                ("depth", _f32p), ("normal", _f32p)]


class OrcCamera(ctypes.Structure):
    _fields_ = [("image_id", ctypes.c_int), ("is_registered", ctypes.c_int),
                ("K", ctypes.c_double * 9), ("R", ctypes.c_double * 9), ("t", ctypes.c_double * 3)]


class OrcPoint(ctypes.Structure):
    _fields_ = [("position", ctypes.c_double * 3), ("image_ids", ctypes.POINTER(ctypes.c_int)),
                ("num_image_ids", ctypes.c_int)]


class OrcViewParams(ctypes.Structure):
    _fields_ = [("ref_image_id", ctypes.c_int), ("num_src", ctypes.c_int),
                ("src_id", ctypes.c_int * MAX_SRC), ("score", ctypes.c_float * MAX_SRC),
                ("baseline_deg", ctypes.c_float * MAX_SRC), ("depth_min", ctypes.c_float),
                ("depth_max", ctypes.c_float)]


class OrcSeed(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("depth", ctypes.c_float),
                ("normal", ctypes.c_float * 3)]


class OrcFuseView(ctypes.Structure):
    _fields_ = [("W", ctypes.c_int), ("H", ctypes.c_int), ("depth", _f32p), ("normal", _f32p),
                ("confidence", _f32p), ("color", _u8p), ("channels", ctypes.c_int),
                ("kinv", ctypes.c_float * 9), ("R", ctypes.c_float * 9), ("t", ctypes.c_float * 3)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_u01.argtypes = [ctypes.c_uint32]
        L.orc_u01.restype = ctypes.c_float
        L.orc_compose.argtypes = [_f64p] * 8
        L.orc_compose.restype = ctypes.c_int
        L.orc_inv3.argtypes = [_f64p, _f64p]
        L.orc_inv3.restype = ctypes.c_int
        L.orc_cost_fp64.argtypes = [ctypes.POINTER(OrcView), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, _u32p, _f64p]
        L.orc_cost_fp64.restype = ctypes.c_double
        L.orc_cost_map_fp64.argtypes = [ctypes.POINTER(OrcView), _f32p, _f32p, _f64p, _u32p,
                                        ctypes.c_int]
        L.orc_cost_list_fp64.argtypes = [ctypes.POINTER(OrcView), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         _f32p, _f32p, _f64p, _u32p, ctypes.c_int]
        L.orc_cost_fp32.argtypes = [ctypes.POINTER(OrcView), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float]
        L.orc_cost_fp32.restype = ctypes.c_float
        L.orc_init_fp32.argtypes = [ctypes.POINTER(OrcView), ctypes.c_uint64, _f32p, _f32p]
        L.orc_halfpass_fp32.argtypes = [ctypes.POINTER(OrcView), ctypes.c_int, ctypes.c_int,
                                        ctypes.c_uint64, _f32p, _f32p, _f32p, ctypes.c_int]
        L.orc_patchmatch_fp32.argtypes = [ctypes.POINTER(OrcView), ctypes.c_int,
                                          ctypes.c_uint64, ctypes.c_int, _f32p, _f32p, _f32p,
                                          ctypes.c_int]
        L.orc_patchmatch_ex.argtypes = [ctypes.POINTER(OrcView), ctypes.c_int, ctypes.c_uint64,
                                        ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p,
                                        ctypes.c_int]
        L.orc_geom_filter_fp32.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _f32p, _f32p,
                                           _f32p, _f32p, _f32p, ctypes.POINTER(OrcGeomSrc),
                                           ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_int]
        L.orc_select_source_views.argtypes = [ctypes.POINTER(OrcCamera), ctypes.c_int,
                                              ctypes.POINTER(OrcPoint), ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(OrcViewParams)]
        L.orc_select_source_views.restype = ctypes.c_int
        L.orc_project_seeds.argtypes = [ctypes.POINTER(OrcCamera), ctypes.POINTER(OrcPoint),
                                        ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                        ctypes.POINTER(OrcSeed)]
        L.orc_project_seeds.restype = ctypes.c_int
        L.orc_apply_seeds.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(OrcSeed),
                                      ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.orc_fuse.argtypes = [ctypes.POINTER(OrcFuseView), ctypes.c_int, ctypes.c_float,
                               _f32p, _f32p, _f32p]
        L.orc_fuse.restype = ctypes.c_longlong
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, _u32p), _p(k, _u32p), _p(out, _u32p))
    return out


def u01(x):
    return lib().orc_u01(int(x))


def inv3(M):
    M = np.ascontiguousarray(M, dtype=np.float64).reshape(9)
    out = np.zeros(9)
    if lib().orc_inv3(_p(M, _f64p), _p(out, _f64p)) != 0:
        raise ValueError("singular matrix")
    return out.reshape(3, 3)


def compose(Kr, Rr, tr, Ks, Rs, ts):
    args = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (Kr, Rr, tr, Ks, Rs, ts)]
    A = np.zeros(9)
    b = np.zeros(3)
    rc = lib().orc_compose(*[_p(a, _f64p) for a in args], _p(A, _f64p), _p(b, _f64p))
    if rc != 0:
        raise ValueError("singular K_ref")
    return A.reshape(3, 3), b


class View:
    """A reference view with its sources, as the oracle sees it.

    ref, srcs: uint8 arrays (H, W) of gray levels (DN, gray = DN / 255), or
    float arrays of gray itself (the reference's float images, nominally in
    [0, 1]; held as float32, as the reference holds them).  Kinv: (3,3)
    float64.  A: list of (3,3), b: list of (3,) float64 -- the FP64-composed
    homography terms.
    """

    def __init__(self, ref, srcs, Kinv, A, b, dmin, dmax, radius=5, stride=1, top_k=3):
        is_float = np.issubdtype(np.asarray(ref).dtype, np.floating)
        assert all(np.issubdtype(np.asarray(s).dtype, np.floating) == is_float for s in srcs)
        if is_float:
            self.gray = [np.ascontiguousarray(ref, dtype=np.float32)] + \
                [np.ascontiguousarray(s, dtype=np.float32) for s in srcs]
            # the 8-bit planes are not read when the float ones are set
            self.ref = np.zeros(self.gray[0].shape, np.uint8)
            self.srcs = [self.ref] * len(srcs)
        else:
            self.gray = None
            self.ref = np.ascontiguousarray(ref, dtype=np.uint8)
            self.srcs = [np.ascontiguousarray(s, dtype=np.uint8) for s in srcs]
        H, W = self.ref.shape
        for s in self.srcs:
            assert s.shape == (H, W)
        assert 1 <= len(self.srcs) <= MAX_SRC
        self.H, self.W = H, W
        v = OrcView()
        v.W, v.H, v.S = W, H, len(self.srcs)
        v.ref = _p(self.ref, _u8p)
        for i, s in enumerate(self.srcs):
            v.src[i] = _p(s, _u8p)
        if self.gray is not None:
            for g in self.gray:
                assert g.shape == (H, W)
            v.ref_gray = _p(self.gray[0], _f32p)
            for i, g in enumerate(self.gray[1:]):
                v.src_gray[i] = _p(g, _f32p)
        Ki = np.asarray(Kinv, dtype=np.float64).reshape(9)
        for i in range(9):
            v.Kinv[i] = Ki[i]
        for s in range(len(self.srcs)):
            a = np.asarray(A[s], dtype=np.float64).reshape(9)
            bb = np.asarray(b[s], dtype=np.float64).reshape(3)
            for i in range(9):
                v.A[s][i] = a[i]
            for i in range(3):
                v.b[s][i] = bb[i]
        v.dmin, v.dmax = float(dmin), float(dmax)
        v.radius, v.stride, v.top_k = int(radius), int(stride), int(top_k)
        self.c = v

    @property
    def ptr(self):
        return ctypes.byref(self.c)

    def cost_fp64(self, x, y, d, n, per_src=False):
        f = ctypes.c_uint32(0)
        ps = np.zeros(MAX_SRC)
        c = lib().orc_cost_fp64(self.ptr, int(x), int(y), float(d), float(n[0]), float(n[1]),
                                float(n[2]), ctypes.byref(f), _p(ps, _f64p))
        if per_src:
            return c, f.value, ps[:self.c.S].copy()
        return c, f.value

    def cost_fp32(self, x, y, d, n):
        return lib().orc_cost_fp32(self.ptr, int(x), int(y), float(d), float(n[0]),
                                   float(n[1]), float(n[2]))

    def cost_map_fp64(self, depth, normal, threads=None):
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        normal = np.ascontiguousarray(normal, dtype=np.float32)
        assert depth.shape == (self.H, self.W) and normal.shape == (self.H, self.W, 3)
        cost = np.zeros((self.H, self.W))
        flags = np.zeros((self.H, self.W), dtype=np.uint32)
        lib().orc_cost_map_fp64(self.ptr, _p(depth, _f32p), _p(normal, _f32p), _p(cost, _f64p),
                                _p(flags, _u32p), threads or os.cpu_count())
        return cost, flags

    def cost_list_fp64(self, xs, ys, depth, normal, threads=None):
        """FP64 cost at listed pixels; depth (n,), normal (n,3) per pixel."""
        xs = np.ascontiguousarray(xs, dtype=np.int32)
        ys = np.ascontiguousarray(ys, dtype=np.int32)
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        normal = np.ascontiguousarray(normal, dtype=np.float32)
        n = len(xs)
        cost = np.zeros(n)
        flags = np.zeros(n, dtype=np.uint32)
        ip = ctypes.POINTER(ctypes.c_int)
        lib().orc_cost_list_fp64(self.ptr, n, xs.ctypes.data_as(ip), ys.ctypes.data_as(ip),
                                 _p(depth, _f32p), _p(normal, _f32p), _p(cost, _f64p),
                                 _p(flags, _u32p), threads or os.cpu_count())
        return cost, flags

    def init_fp32(self, seed):
        depth = np.zeros((self.H, self.W), dtype=np.float32)
        normal = np.zeros((self.H, self.W, 3), dtype=np.float32)
        lib().orc_init_fp32(self.ptr, int(seed), _p(depth, _f32p), _p(normal, _f32p))
        return depth, normal

    def halfpass_fp32(self, iteration, parity, seed, depth, normal, cost, threads=None):
        for a in (depth, normal, cost):
            assert a.dtype == np.float32 and a.flags.c_contiguous
        lib().orc_halfpass_fp32(self.ptr, int(iteration), int(parity), int(seed),
                                _p(depth, _f32p), _p(normal, _f32p), _p(cost, _f32p),
                                threads or os.cpu_count())

    def patchmatch_fp32(self, iterations, seed, threads=None, init=None, cost_mode=COST_FP32):
        """Full run; cost_mode COST_FP64 evaluates every cost in double."""
        depth = np.zeros((self.H, self.W), dtype=np.float32)
        normal = np.zeros((self.H, self.W, 3), dtype=np.float32)
        cost = np.zeros((self.H, self.W), dtype=np.float32)
        given = 0
        if init is not None:
            depth[...] = init[0]
            normal[...] = init[1]
            given = 1
        lib().orc_patchmatch_ex(self.ptr, int(iterations), int(seed), given, int(cost_mode),
                                _p(depth, _f32p), _p(normal, _f32p), _p(cost, _f32p),
                                threads or os.cpu_count())
        return depth, normal, cost


def geom_filter_fp32(depth, normal, confidence, Kinv, R, t, srcs, depth_tol=0.12,
                     normal_tol_cos=-1.0, min_consistent=2):
    """In-place geometric consistency filter.  srcs: list of (K, R, t, depth, normal)."""
    H, W = depth.shape
    arr = (OrcGeomSrc * max(1, len(srcs)))()
    keep = []
    for i, (K, Rs, ts, d, n) in enumerate(srcs):
        d = np.ascontiguousarray(d, dtype=np.float32)
        n = np.ascontiguousarray(n, dtype=np.float32)
        keep += [d, n]
        for j, val in enumerate(np.asarray(K, dtype=np.float32).reshape(9)):
            arr[i].K[j] = val
        for j, val in enumerate(np.asarray(Rs, dtype=np.float32).reshape(9)):
            arr[i].R[j] = val
        for j, val in enumerate(np.asarray(ts, dtype=np.float32).reshape(3)):
            arr[i].t[j] = val
        arr[i].depth = _p(d, _f32p)
        arr[i].normal = _p(n, _f32p)
    ki = np.ascontiguousarray(Kinv, dtype=np.float32).reshape(9)
    Rr = np.ascontiguousarray(R, dtype=np.float32).reshape(9)
    tt = np.ascontiguousarray(t, dtype=np.float32).reshape(3)
    lib().orc_geom_filter_fp32(W, H, _p(depth, _f32p), _p(normal, _f32p), _p(confidence, _f32p),
                               _p(ki, _f32p), _p(Rr, _f32p), _p(tt, _f32p), arr, len(srcs),
                               float(depth_tol), float(normal_tol_cos), int(min_consistent))


class SfM:
    """Cameras (dicts with K, R, t, optional registered) and points
    ((position, image_ids) pairs) packed for the restatement."""

    def __init__(self, cameras, points):
        self.cams = (OrcCamera * max(1, len(cameras)))()
        for i, c in enumerate(cameras):
            cc = self.cams[i]
            cc.image_id = i
            cc.is_registered = int(c.get("registered", True))
            for name in ("K", "R", "t"):
                for j, v in enumerate(np.asarray(c[name], np.float64).reshape(-1)):
                    getattr(cc, name)[j] = v
        self.n_cams = len(cameras)
        self.pts = (OrcPoint * max(1, len(points)))()
        self._ids = []
        for k, (pos, ids) in enumerate(points):
            ids = np.ascontiguousarray(ids, dtype=np.int32)
            self._ids.append(ids)
            for j in range(3):
                self.pts[k].position[j] = float(pos[j])
            self.pts[k].image_ids = ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
            self.pts[k].num_image_ids = len(ids)
        self.n_pts = len(points)

    def select_source_views(self, num_sources=5):
        out = (OrcViewParams * max(1, self.n_cams))()
        n = lib().orc_select_source_views(self.cams, self.n_cams, self.pts, self.n_pts,
                                          int(num_sources), out)
        return [{"ref_image_id": out[i].ref_image_id,
                 "sources": [(out[i].src_id[s], out[i].score[s], out[i].baseline_deg[s])
                             for s in range(out[i].num_src)],
                 "depth_min": out[i].depth_min, "depth_max": out[i].depth_max}
                for i in range(n)]

    def seeds(self, ref_id, dmin, dmax):
        """[(x, y, depth, (nx, ny, nz))] in point order."""
        out = (OrcSeed * max(1, self.n_pts))()
        n = lib().orc_project_seeds(ctypes.byref(self.cams[ref_id]), self.pts, self.n_pts,
                                    float(dmin), float(dmax), out)
        return [(out[i].x, out[i].y, out[i].depth, tuple(out[i].normal)) for i in range(n)]


def apply_seeds(depth, normal, seeds, radius=50):
    """Paint seeds [(x, y, depth, n)] in order over depth (H,W) / normal (H,W,3), in place."""
    H, W = depth.shape
    arr = (OrcSeed * max(1, len(seeds)))()
    for i, (x, y, d, n) in enumerate(seeds):
        arr[i].x, arr[i].y, arr[i].depth = int(x), int(y), float(d)
        for j in range(3):
            arr[i].normal[j] = float(n[j])
    assert depth.dtype == np.float32 and normal.dtype == np.float32
    lib().orc_apply_seeds(W, H, arr, len(seeds), int(radius), _p(depth, _f32p), _p(normal, _f32p))


def fuse(maps, min_confidence=0.0):
    """Restatement of the fusion; maps as for the product's Context.fuse.
    K^-1 is inverted in FP64 and rounded to FP32, as are R and t."""
    n = len(maps)
    arr = (OrcFuseView * n)()
    keep = []
    total = 0
    for i, m in enumerate(maps):
        g = arr[i]
        d = np.ascontiguousarray(m["depth"], np.float32)
        keep.append(d)
        g.H, g.W = d.shape
        total += d.size
        g.depth = _p(d, _f32p)
        for k in ("normal", "confidence"):
            if m.get(k) is not None:
                a = np.ascontiguousarray(m[k], np.float32)
                keep.append(a)
                setattr(g, k, _p(a, _f32p))
        if m.get("color") is not None:
            c = np.ascontiguousarray(m["color"], np.uint8)
            keep.append(c)
            g.color = c.ctypes.data_as(_u8p)
            g.channels = 1 if c.ndim == 2 else c.shape[2]
        ki = inv3(np.asarray(m["K"], np.float64))
        for j, val in enumerate(np.asarray(ki, np.float64).reshape(9)):
            g.kinv[j] = float(np.float32(val))
        for j, val in enumerate(np.asarray(m["R"], np.float64).reshape(9)):
            g.R[j] = float(np.float32(val))
        for j, val in enumerate(np.asarray(m["t"], np.float64).reshape(3)):
            g.t[j] = float(np.float32(val))
    pos = np.zeros((total, 3), np.float32)
    nrm = np.zeros((total, 3), np.float32)
    col = np.zeros((total, 3), np.float32)
    N = lib().orc_fuse(arr, n, float(min_confidence), _p(pos, _f32p), _p(nrm, _f32p),
                       _p(col, _f32p))
    return pos[:N], nrm[:N], col[:N]
