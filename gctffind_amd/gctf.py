"""ctypes binding of libgctf_hip.so (include/gctf_hip.h).

The classes mirror the reference's C++ call surface (FindCTF/CFindCTFInc.h):
CRescaleImage, CCTFTheory and CFindCtf2D, with the same method names and
argument meaning, so the parity tests read like calls into the reference.
Device buffers are passed as raw addresses; torch is only used by callers
to own device memory.

The library must be built (python -m gctffind_amd.build); loading fails
loudly when it is missing.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgctf_hip.so")

GCTF_OK, GCTF_ERR_ARG, GCTF_ERR_HIP, GCTF_ERR_STATE, GCTF_ERR_UNSUPPORTED, GCTF_ERR_NOMEM = range(6)
DTYPE_F32, DTYPE_F16 = 0, 1

EXPORTS = [
    "gctf_create", "gctf_destroy", "gctf_error_string", "gctf_set_stream", "gctf_synchronize",
    "gctf_host_alloc", "gctf_host_free",
    "gctf_rescale_setup", "gctf_rescale", "gctf_rescale_prefetch", "gctf_scaled_image", "gctf_setup1", "gctf_setup2",
    "gctf_set_phase", "gctf_set_search1d", "gctf_gen_half_spectrum", "gctf_do1d", "gctf_refine1d", "gctf_do2d",
    "gctf_do2d_batch_static",
    "gctf_refine", "gctf_spect_size", "gctf_get_half_spect", "gctf_set_half_spect",
    "gctf_full_spectrum", "gctf_tile_grid", "gctf_ring_pixels", "gctf_radial_average", "gctf_estimate_batch", "gctf_tune", "gctf_profile",
    "gctf_profile_report", "gctf_grid_search_batch",
    "gctf_estimate_list", "gctf_estimate_series", "gctf_autotune",
    "gctf_learned_program_default", "gctf_radial_batch", "gctf_set_proposal_params",
    "gctf_estimate_batch_learned",
    "gctf_net_param_count", "gctf_net_create", "gctf_net_destroy", "gctf_net_error_string",
    "gctf_net_set_params", "gctf_net_get_params", "gctf_net_reset_optimizer", "gctf_net_forward",
    "gctf_net_loss_grad", "gctf_net_adam_step", "gctf_net_synchronize",
]


class GctfParams(ctypes.Structure):
    _fields_ = [("kv", ctypes.c_float), ("cs_mm", ctypes.c_float),
                ("amp_contrast", ctypes.c_float), ("ext_phase_deg", ctypes.c_float),
                ("ext_phase_range_deg", ctypes.c_float), ("tile_size", ctypes.c_int),
                ("log_spect", ctypes.c_int)]


class GctfResult(ctypes.Structure):
    _fields_ = [("df_min", ctypes.c_float), ("df_max", ctypes.c_float),
                ("azimuth_deg", ctypes.c_float), ("ext_phase_deg", ctypes.c_float),
                ("score", ctypes.c_float), ("ctf_res", ctypes.c_float)]

    def as_dict(self):
        return dict(df_min=self.df_min, df_max=self.df_max, azimuth=self.azimuth_deg,
                    ext_phase=self.ext_phase_deg, score=self.score, ctf_res=self.ctf_res)


class GctfGrid(ctypes.Structure):
    _fields_ = [("df_lo", ctypes.c_float), ("df_hi", ctypes.c_float), ("n_df", ctypes.c_int),
                ("ast_lo", ctypes.c_float), ("ast_hi", ctypes.c_float), ("n_ast", ctypes.c_int),
                ("az_lo", ctypes.c_float), ("az_hi", ctypes.c_float), ("n_az", ctypes.c_int),
                ("res_lo", ctypes.c_float), ("res_hi", ctypes.c_float)]


# the candidate grid of the 256-micrograph BF16 search (256 df x 32 ast x 36 az,
# ring 1/30 - 1/4 1/A)
GRID_256x32x36 = dict(df_lo=3000.0, df_hi=43000.0, n_df=256, ast_lo=0.0, ast_hi=0.06, n_ast=32,
                      az_lo=-90.0, az_hi=90.0, n_az=36, res_lo=30.0, res_hi=4.0)


class GctfImage(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("pix", ctypes.c_float)]


class GctfSearchProgram(ctypes.Structure):
    _fields_ = [("n_ast", ctypes.c_int), ("n_ang", ctypes.c_int), ("n_mag", ctypes.c_int),
                ("n_angl", ctypes.c_int), ("n_df", ctypes.c_int), ("iters", ctypes.c_int),
                ("df_minstep", ctypes.c_float), ("fine_n", ctypes.c_int), ("fine_step", ctypes.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GctfProposal(ctypes.Structure):
    _fields_ = [("df_mean", ctypes.c_float), ("astig", ctypes.c_float), ("azimuth_deg", ctypes.c_float),
                ("confidence", ctypes.c_float), ("fallback", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GctfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gctf error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libgctf_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gctffind_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cf, fp = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_void_p
    sig = {
        "gctf_create": (ci, [ctypes.POINTER(GctfParams), ci, ctypes.POINTER(vp)]),
        "gctf_destroy": (None, [vp]),
        "gctf_error_string": (ctypes.c_char_p, [vp]),
        "gctf_set_stream": (ci, [vp, vp]),
        "gctf_synchronize": (ci, [vp]),
        "gctf_host_alloc": (ci, [ctypes.c_size_t, ctypes.POINTER(vp)]),
        "gctf_host_free": (None, [vp]),
        "gctf_rescale_setup": (ci, [vp, ci, ci, cf, fp, fp]),
        "gctf_rescale": (ci, [vp, fp, ci]),
        "gctf_rescale_prefetch": (ci, [vp, fp, ci, ci]),
        "gctf_scaled_image": (ci, [vp, ctypes.POINTER(vp)]),
        "gctf_setup1": (ci, [vp, cf, cf, cf, cf, cf, cf]),
        "gctf_setup2": (ci, [vp, ci, ci]),
        "gctf_set_phase": (ci, [vp, cf, cf]),
        "gctf_set_search1d": (ci, [vp, cf, cf, cf]),
        "gctf_gen_half_spectrum": (ci, [vp, fp]),
        "gctf_do1d": (ci, [vp, ctypes.POINTER(GctfResult)]),
        "gctf_refine1d": (ci, [vp, cf, cf, ctypes.POINTER(GctfResult)]),
        "gctf_do2d": (ci, [vp, ctypes.POINTER(GctfResult)]),
        "gctf_refine": (ci, [vp, fp, fp, fp, fp, ctypes.POINTER(GctfResult)]),
        "gctf_spect_size": (ci, [vp, fp, ci]),
        "gctf_get_half_spect": (ci, [vp, ci, fp]),
        "gctf_set_half_spect": (ci, [vp, fp]),
        "gctf_full_spectrum": (ci, [vp, fp]),
        "gctf_tile_grid": (ci, [ci, ci, ci, fp]),
        "gctf_ring_pixels": (ci, [vp, cf, cf, cf, fp, ci, fp]),
        "gctf_radial_average": (ci, [vp, vp, ci, vp, ci]),
        "gctf_estimate_batch": (ci, [vp, vp, ci, ci, ci, ci, ctypes.c_longlong, cf, fp]),
        "gctf_tune": (ci, [vp, ci, ci, ci, ci, ci]),
        "gctf_profile": (ci, [vp, ci]),
        "gctf_profile_report": (ci, [vp, ci, fp, fp, fp, fp]),
        "gctf_grid_search_batch": (ci, [vp, vp, ci, ci, ci, ci, ctypes.c_longlong, cf,
                                        ctypes.POINTER(GctfGrid), fp, fp, fp, fp]),
        "gctf_do2d_batch_static": (ci, [vp, vp, ci, vp]),
        "gctf_estimate_list": (ci, [vp, ctypes.POINTER(GctfImage), ci, fp]),
        "gctf_estimate_series": (ci, [vp, vp, ci, ci, ci, ci, ctypes.c_longlong, cf, fp, ci, fp]),
        "gctf_autotune": (ci, [vp, vp, ci, ci, ci, ci, ctypes.c_longlong, cf, fp, fp, fp, fp]),
        "gctf_learned_program_default": (ci, [ctypes.POINTER(GctfSearchProgram)]),
        "gctf_radial_batch": (ci, [vp, vp, ci, ci, ci, ci, ctypes.c_longlong, cf, vp, ci]),
        "gctf_set_proposal_params": (ci, [vp, fp, ctypes.c_longlong]),
        "gctf_estimate_batch_learned": (ci, [vp, vp, ci, ci, ci, ci, ctypes.c_longlong, cf,
                                             ctypes.POINTER(GctfSearchProgram), cf, cf, fp, fp]),
        "gctf_net_param_count": (ctypes.c_longlong, []),
        "gctf_net_create": (ci, [ci, ci, ctypes.POINTER(vp)]),
        "gctf_net_destroy": (None, [vp]),
        "gctf_net_error_string": (ctypes.c_char_p, [vp]),
        "gctf_net_set_params": (ci, [vp, fp, ctypes.c_longlong]),
        "gctf_net_get_params": (ci, [vp, fp, ctypes.c_longlong]),
        "gctf_net_reset_optimizer": (ci, [vp]),
        "gctf_net_forward": (ci, [vp, vp, ci, ci, ci, fp]),
        "gctf_net_loss_grad": (ci, [vp, vp, ci, vp, ci, ci, fp, fp]),
        "gctf_net_adam_step": (ci, [vp, cf, cf, cf, cf, cf]),
        "gctf_net_synchronize": (ci, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One gctf_ctx: a GPU, a stream and the CFindCtf2D-shaped state."""

    def __init__(self, kv=300.0, cs=2.7, ac=0.07, ext_phase=(0.0, 0.0), tile=512, log_spect=0,
                 device=0):
        self.L = load()
        self.params = GctfParams(kv, cs, ac, ext_phase[0], ext_phase[1], tile, log_spect)
        h = ctypes.c_void_p()
        rc = self.L.gctf_create(ctypes.byref(self.params), device, ctypes.byref(h))
        if rc != GCTF_OK:
            raise GctfError(rc, "gctf_create failed")
        self.h = h
        self.tile = tile

    def close(self):
        if getattr(self, "h", None):
            self.L.gctf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != GCTF_OK:
            msg = self.L.gctf_error_string(self.h)
            raise GctfError(rc, msg.decode() if msg else "")
        return rc

    # ---- CRescaleImage ----------------------------------------------------
    def rescale_setup(self, nx, ny, pix):
        size = np.zeros(3, np.int32)
        pixn = np.zeros(1, np.float32)
        self.check(self.L.gctf_rescale_setup(self.h, nx, ny, pix, _np_ptr(size), _np_ptr(pixn)))
        return int(size[0]), int(size[1]), int(size[2]), float(pixn[0])

    def rescale(self, img, on_device=False):
        if on_device:
            self.check(self.L.gctf_rescale(self.h, ctypes.c_void_p(img), 1))
        else:
            img = np.ascontiguousarray(img, np.float32)
            self.check(self.L.gctf_rescale(self.h, _np_ptr(img), 0))

    def rescale_prefetch(self, img):
        """Starts the device copy of host image img, which the next rescale
        of the same array uses.  img must be a C-contiguous float32 array
        that the caller keeps, unchanged, until that rescale returns."""
        if not (isinstance(img, np.ndarray) and img.dtype == np.float32 and img.flags["C_CONTIGUOUS"]
                and img.ndim == 2):
            raise GctfError(GCTF_ERR_ARG, "prefetch needs a 2-D C-contiguous float32 array")
        self.check(self.L.gctf_rescale_prefetch(self.h, _np_ptr(img), img.shape[1], img.shape[0]))

    def scaled_image(self):
        p = ctypes.c_void_p()
        self.check(self.L.gctf_scaled_image(self.h, ctypes.byref(p)))
        return p.value

    # ---- CCTFTheory + CFindCtf2D ----------------------------------------
    def setup1(self, kv, cs, ac, pix, ast_tol=100.0, ext_phase_rad=0.0):
        self.check(self.L.gctf_setup1(self.h, kv, cs, ac, pix, ast_tol, ext_phase_rad))

    def setup2(self, nx, ny):
        self.check(self.L.gctf_setup2(self.h, nx, ny))

    def set_phase(self, p0, rng):
        self.check(self.L.gctf_set_phase(self.h, p0, rng))

    def set_search1d(self, df_lo=0.0, df_hi=0.0, step=0.0):
        """Initial Do1D defocus scan [df_lo, df_hi] A in steps of `step` A
        (all 0: the reference's 3000 p^2 .. 30000 p^2 in 500 steps)."""
        self.check(self.L.gctf_set_search1d(self.h, df_lo, df_hi, step))

    def gen_half_spectrum(self, dev_pad_img):
        self.check(self.L.gctf_gen_half_spectrum(self.h, ctypes.c_void_p(dev_pad_img)))

    def do1d(self):
        r = GctfResult()
        self.check(self.L.gctf_do1d(self.h, ctypes.byref(r)))
        return r.as_dict()

    def refine1d(self, df, rng):
        r = GctfResult()
        self.check(self.L.gctf_refine1d(self.h, df, rng, ctypes.byref(r)))
        return r.as_dict()

    def do2d(self):
        r = GctfResult()
        self.check(self.L.gctf_do2d(self.h, ctypes.byref(r)))
        return r.as_dict()

    def do2d_batch_static(self, dev_spect_ptr, n, dev_out_ptr):
        """Do1D + Do2D on n <= 16 device half spectra, results into n device
        gctf_result rows; allocates nothing (static scratch per GPU).
        Enqueued only: synchronize() before reading the results."""
        self.check(self.L.gctf_do2d_batch_static(self.h, ctypes.c_void_p(dev_spect_ptr), n,
                                                 ctypes.c_void_p(dev_out_ptr)))

    def refine(self, df_mean, ast_ratio, ast_angle, ext_phase):
        a = [np.asarray(v, np.float32) for v in (df_mean, ast_ratio, ast_angle, ext_phase)]
        r = GctfResult()
        self.check(self.L.gctf_refine(self.h, *[_np_ptr(v) for v in a], ctypes.byref(r)))
        return r.as_dict()

    def spect_size(self, half=True):
        s = np.zeros(2, np.int32)
        self.check(self.L.gctf_spect_size(self.h, _np_ptr(s), int(half)))
        return int(s[0]), int(s[1])

    def get_half_spect(self, raw=False):
        out = np.zeros((self.tile, self.tile // 2 + 1), np.float32)
        self.check(self.L.gctf_get_half_spect(self.h, int(raw), _np_ptr(out)))
        return out

    def set_half_spect(self, spect):
        spect = np.ascontiguousarray(spect, np.float32)
        self.check(self.L.gctf_set_half_spect(self.h, _np_ptr(spect)))

    def full_spectrum(self):
        out = np.zeros((self.tile, self.tile), np.float32)
        self.check(self.L.gctf_full_spectrum(self.h, _np_ptr(out)))
        return out

    # ---- index tables ----------------------------------------------------
    def ring_pixels(self, pix, res_lo=0.0, res_hi=0.0):
        """Half-spectrum indices of the 2D correlation ring at pixel size pix."""
        n = ctypes.c_int()
        self.check(self.L.gctf_ring_pixels(self.h, pix, res_lo, res_hi, None, 0,
                                           ctypes.cast(ctypes.pointer(n), ctypes.c_void_p)))
        idx = np.zeros(n.value, np.int32)
        self.check(self.L.gctf_ring_pixels(self.h, pix, res_lo, res_hi, _np_ptr(idx), idx.size,
                                           ctypes.cast(ctypes.pointer(n), ctypes.c_void_p)))
        return idx

    def radial_average(self, dev_spect_ptr, n, dev_out_ptr, ld=None):
        """GRadialAvg on n device half spectra into device rows of ld floats."""
        ld = self.tile // 2 + 1 if ld is None else ld
        self.check(self.L.gctf_radial_average(self.h, ctypes.c_void_p(dev_spect_ptr), n,
                                              ctypes.c_void_p(dev_out_ptr), ld))

    # ---- batched path ----------------------------------------------------
    def estimate_batch(self, dev_ptr, n, nx, ny, pix, dtype=DTYPE_F32, stride=None):
        stride = nx * ny if stride is None else stride
        res = (GctfResult * n)()
        self.check(self.L.gctf_estimate_batch(self.h, ctypes.c_void_p(dev_ptr), dtype, n, nx, ny,
                                              stride, pix, ctypes.cast(res, ctypes.c_void_p)))
        return [r.as_dict() for r in res]

    def grid_search_batch(self, dev_ptr, n, nx, ny, pix, grid, dtype=DTYPE_F32, stride=None,
                          debug=False):
        stride = nx * ny if stride is None else stride
        g = GctfGrid(**grid)
        m = g.n_df * g.n_ast * g.n_az
        res = (GctfResult * n)()
        cc = np.zeros((n, m), np.float32) if debug else None
        top = np.zeros((n, 16), np.int32) if debug else None
        rs = np.zeros((n, 16), np.float32) if debug else None
        p = (lambda a: _np_ptr(a) if a is not None else None)
        self.check(self.L.gctf_grid_search_batch(self.h, ctypes.c_void_p(dev_ptr), dtype, n, nx, ny, stride,
                                                 pix, ctypes.byref(g), ctypes.cast(res, ctypes.c_void_p),
                                                 p(cc), p(top), p(rs)))
        out = [r.as_dict() for r in res]
        return (out, cc, top, rs) if debug else out

    def estimate_list(self, images):
        """images: sequence of dicts (data: device address, dtype, nx, ny, pix)."""
        n = len(images)
        arr = (GctfImage * n)(*[GctfImage(ctypes.c_void_p(im["data"]), im.get("dtype", DTYPE_F32), im["nx"],
                                          im["ny"], im["pix"]) for im in images])
        res = (GctfResult * n)()
        self.check(self.L.gctf_estimate_list(self.h, arr, n, ctypes.cast(res, ctypes.c_void_p)))
        return [r.as_dict() for r in res]

    def estimate_series(self, dev_ptr, n, nx, ny, pix, tilts, tomo=True, dtype=DTYPE_F32, stride=None):
        stride = nx * ny if stride is None else stride
        t = np.ascontiguousarray(tilts, dtype=np.float32)
        res = (GctfResult * n)()
        self.check(self.L.gctf_estimate_series(self.h, ctypes.c_void_p(dev_ptr), dtype, n, nx, ny, stride, pix,
                                               _np_ptr(t), int(bool(tomo)), ctypes.cast(res, ctypes.c_void_p)))
        return [r.as_dict() for r in res]

    def autotune(self, dev_ptr, n, nx, ny, pix, dtype=DTYPE_F32, stride=None):
        stride = nx * ny if stride is None else stride
        c, sp, ms, cb = ctypes.c_int(), ctypes.c_int(), ctypes.c_float(), ctypes.c_int()
        self.check(self.L.gctf_autotune(self.h, ctypes.c_void_p(dev_ptr), dtype, n, nx, ny, stride, pix,
                                        ctypes.cast(ctypes.pointer(c), ctypes.c_void_p),
                                        ctypes.cast(ctypes.pointer(sp), ctypes.c_void_p),
                                        ctypes.cast(ctypes.pointer(ms), ctypes.c_void_p),
                                        ctypes.cast(ctypes.pointer(cb), ctypes.c_void_p)))
        return {"chunk": c.value, "tile_splits": sp.value, "ms": ms.value, "search_block": cb.value}

    def radial_batch(self, dev_ptr, n, nx, ny, pix, dev_out_ptr, ld, dtype=DTYPE_F32, stride=None):
        """Spectrum stages + radial average into a device buffer [n][ld]."""
        stride = nx * ny if stride is None else stride
        self.check(self.L.gctf_radial_batch(self.h, ctypes.c_void_p(dev_ptr), dtype, n, nx, ny, stride, pix,
                                            ctypes.c_void_p(dev_out_ptr), ld))

    def set_proposal_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self.check(self.L.gctf_set_proposal_params(self.h, _np_ptr(p), p.size))

    def estimate_batch_learned(self, dev_ptr, n, nx, ny, pix, program=None, conf_threshold=0.8,
                               window=300.0, dtype=DTYPE_F32, stride=None, proposals=False):
        stride = nx * ny if stride is None else stride
        prog = None if program is None else ctypes.byref(GctfSearchProgram(**program))
        res = (GctfResult * n)()
        props = (GctfProposal * n)()
        self.check(self.L.gctf_estimate_batch_learned(
            self.h, ctypes.c_void_p(dev_ptr), dtype, n, nx, ny, stride, pix, prog, conf_threshold, window,
            ctypes.cast(res, ctypes.c_void_p), ctypes.cast(props, ctypes.c_void_p)))
        out = [r.as_dict() for r in res]
        return (out, [p.as_dict() for p in props]) if proposals else out

    def tune(self, chunk=0, tile_splits=0, use_graph=0, search_block=0, list_queue=0):
        self.check(self.L.gctf_tune(self.h, chunk, tile_splits, use_graph, search_block, list_queue))

    def profile(self, enable=True):
        self.check(self.L.gctf_profile(self.h, int(enable)))

    def profile_report(self, max_rows=64):
        names = ctypes.create_string_buffer(64 * max_rows)
        tot = np.zeros(max_rows, np.float64)
        cnt = np.zeros(max_rows, np.int32)
        n = ctypes.c_int()
        self.check(self.L.gctf_profile_report(self.h, max_rows, ctypes.cast(names, ctypes.c_void_p),
                                              _np_ptr(tot), _np_ptr(cnt), ctypes.cast(ctypes.pointer(n), ctypes.c_void_p)))
        out = {}
        for i in range(n.value):
            nm = names.raw[64 * i: 64 * (i + 1)].split(b"\0")[0].decode()
            out[nm] = (float(tot[i]), int(cnt[i]))
        return out

    def synchronize(self):
        self.check(self.L.gctf_synchronize(self.h))


def tile_grid(nx, ny, tile=512):
    """(tiles x, tiles y, offset x, offset y, overlap, step) of CGenAvgSpectrum::SetSizes."""
    out = np.zeros(6, np.int32)
    rc = load().gctf_tile_grid(nx, ny, tile, _np_ptr(out))
    if rc != GCTF_OK:
        raise GctfError(rc, "gctf_tile_grid")
    return tuple(int(v) for v in out)


def learned_program_default():
    L = load()
    p = GctfSearchProgram()
    L.gctf_learned_program_default(ctypes.byref(p))
    return p.as_dict()


class Net:
    """Trainer of the proposal network (gctf_net_* in the C-ABI)."""

    def __init__(self, max_batch, device=0):
        self.L = load()
        self.h = ctypes.c_void_p()
        rc = self.L.gctf_net_create(device, max_batch, ctypes.byref(self.h))
        if rc != GCTF_OK:
            raise GctfError(rc, "gctf_net_create failed")
        self.n_params = int(self.L.gctf_net_param_count())

    def close(self):
        if self.h:
            self.L.gctf_net_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != GCTF_OK:
            raise GctfError(rc, self.L.gctf_net_error_string(self.h).decode())

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self.check(self.L.gctf_net_set_params(self.h, _np_ptr(p), p.size))

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        self.check(self.L.gctf_net_get_params(self.h, _np_ptr(p), p.size))
        return p

    def reset_optimizer(self):
        self.check(self.L.gctf_net_reset_optimizer(self.h))

    def forward(self, dev_radial_ptr, ld, n, bf16=True):
        out = np.zeros((n, 4), np.float32)
        self.check(self.L.gctf_net_forward(self.h, ctypes.c_void_p(dev_radial_ptr), ld, n, int(bf16),
                                           _np_ptr(out)))
        return out

    def loss_grad(self, dev_radial_ptr, ld, dev_target_ptr, n, bf16=True, want_loss=True, want_grad=False):
        loss = np.zeros(1, np.float64)
        g = np.zeros(self.n_params, np.float32) if want_grad else None
        self.check(self.L.gctf_net_loss_grad(self.h, ctypes.c_void_p(dev_radial_ptr), ld,
                                             ctypes.c_void_p(dev_target_ptr), n, int(bf16),
                                             _np_ptr(loss) if want_loss else None,
                                             _np_ptr(g) if want_grad else None))
        return (float(loss[0]) if want_loss else None), g

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0):
        self.check(self.L.gctf_net_adam_step(self.h, lr, beta1, beta2, eps, weight_decay))

    def synchronize(self):
        self.check(self.L.gctf_net_synchronize(self.h))


class CFindCtf2D:
    """CFindCtf2D-shaped facade (CFindCTFInc.h:541-615) over a Context."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.m_fDfMin = self.m_fDfMax = self.m_fAstAng = self.m_fExtPhase = self.m_fScore = 0.0

    def _take(self, r):
        self.m_fDfMin, self.m_fDfMax = r["df_min"], r["df_max"]
        self.m_fAstAng, self.m_fExtPhase, self.m_fScore = r["azimuth"], r["ext_phase"], r["score"]
        return r

    def Setup1(self, kv, cs, ac, pix, ast_tol=100.0, ext_phase_rad=0.0):
        self.ctx.setup1(kv, cs, ac, pix, ast_tol, ext_phase_rad)

    def Setup2(self, size):
        self.ctx.setup2(int(size[0]), int(size[1]))

    def SetPhase(self, p0, rng):
        self.ctx.set_phase(p0, rng)

    def GenHalfSpectrum(self, dev_pad_img):
        self.ctx.gen_half_spectrum(dev_pad_img)

    def Do1D(self):
        return self._take(self.ctx.do1d())

    def Refine1D(self, df, rng):
        return self._take(self.ctx.refine1d(df, rng))

    def Do2D(self):
        return self._take(self.ctx.do2d())

    def Refine(self, df_mean, ast_ratio, ast_angle, ext_phase):
        return self._take(self.ctx.refine(df_mean, ast_ratio, ast_angle, ext_phase))

    def GetHalfSpect(self, raw=False):
        return self.ctx.get_half_spect(raw)

    def SetHalfSpect(self, spect):
        self.ctx.set_half_spect(spect)

    def GetSpectSize(self, half=True):
        return self.ctx.spect_size(half)

    def GenFullSpectrum(self):
        return self.ctx.full_spectrum()
