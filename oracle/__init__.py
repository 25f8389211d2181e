"""ctypes binding of the CPU oracle (oracle/gctf_oracle.c, oracle/net_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this module as the checker.  The product package
(gctffind_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_int, _c_float, _c_void = ctypes.c_int, ctypes.c_float, ctypes.c_void_p

# stage ids of the candidate trace (gctf_oracle.c ORC_ST_*)
ST_1D, ST_1D_REFINE, ST_INIT, ST_ASTIG, ST_ASTMAG, ST_ASTANG, ST_DFMEAN, ST_PHASE = range(8)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    sig = {
        "orc_fft": (None, [_f64p, _c_int, _c_int]),
        "orc_wavelength": (_c_float, [_c_float]),
        "orc_ctf_params": (None, [_c_float] * 5 + [_f32p]),
        "orc_rescale_sizes": (None, [_c_int, _c_int, _c_float, _i32p, _f32p]),
        "orc_rescale": (None, [_f32p, _c_int, _c_int, _c_float, _f64p]),
        "orc_tile_grid": (None, [_c_int, _c_int, _c_int, _i32p]),
        "orc_round_edge_table": (None, [_c_int, _f64p]),
        "orc_avg_spectrum": (None, [_f64p, _c_int, _c_int, _c_int, _c_int, _f64p]),
        "orc_rm_background": (None, [_f64p, _f64p, _c_int, _c_int, _c_int, _c_float]),
        "orc_moments": (None, [_f64p, _c_int, _f64p]),
        "orc_gate": (_c_int, [_f64p, _c_int]),
        "orc_lowpass": (None, [_f64p, _c_int, _c_int]),
        "orc_radial_table": (_c_int, [_c_int, _c_int, _c_int, _c_void, _c_void]),
        "orc_radial_table_f64": (_c_int, [_c_int, _c_int, _c_void, _c_void, _c_void]),
        "orc_radial_avg": (None, [_f64p, _c_int, _c_int, _f64p]),
        "orc_ctf1d": (None, [_f32p, _c_float, _c_float, _c_int, _f64p]),
        "orc_cc1d": (_c_float, [_f64p, _f64p, _c_int, _c_float, _c_float, _c_float]),
        "orc_ctf2d": (None, [_f32p, _c_float, _c_float, _c_float, _c_float, _c_int, _c_int, _f64p]),
        "orc_ring_pixels": (_c_int, [_c_int, _c_int, _c_float, _c_float, _c_void]),
        "orc_cc2d": (ctypes.c_double, [_f64p, _f64p, _c_int, _c_int, _c_float, _c_float, _c_float]),
        "orc_spectral_cc": (_c_int, [_f64p, _f64p, _c_int, _c_int, _c_int, _c_void]),
        "orc_net_param_count": (ctypes.c_long, []),
        "orc_net_forward": (None, [_f64p, _f32p, _c_int, _c_int, _f64p]),
        "orc_net_loss_grad": (ctypes.c_double, [_f64p, _f32p, _c_int, _f32p, _c_int, _c_void]),
        "orc_fc_create": (_c_void, []),
        "orc_fc_destroy": (None, [_c_void]),
        "orc_fc_setup": (None, [_c_void, _c_float, _c_float, _c_float, _c_float, _c_int, _c_int, _c_float]),
        "orc_fc_set_search1d": (None, [_c_void, _c_float, _c_float, _c_float]),
        "orc_fc_trace_enable": (None, [_c_void, _c_int]),
        "orc_fc_trace_count": (_c_int, [_c_void]),
        "orc_fc_trace_get": (None, [_c_void, _f32p]),
        "orc_fc_gen_half_spectrum": (None, [_c_void, _f64p, _c_int, _c_int, _c_int]),
        "orc_fc_set_half_spect": (None, [_c_void, _f64p]),
        "orc_fc_set_raw_spect": (None, [_c_void, _f64p]),
        "orc_fc_get_spects": (None, [_c_void, _c_void, _c_void, _c_void]),
        "orc_fc_lowpassed": (_c_int, [_c_void]),
        "orc_fc_set_phase": (None, [_c_void, _c_float, _c_float]),
        "orc_fc_do1d": (None, [_c_void]),
        "orc_fc_refine1d": (None, [_c_void, _c_float, _c_float]),
        "orc_fc_do2d": (None, [_c_void]),
        "orc_fc_refine": (None, [_c_void, _f32p, _f32p, _f32p, _f32p]),
        "orc_fc_results": (None, [_c_void, _f32p]),
        "orc_fc_full_spectrum": (None, [_c_void, _f64p]),
        "orc_grid_cc": (None, [_f64p, _c_int, _c_int, _c_float, _c_float, _c_float, _c_float, _c_float,
                               _f32p, _i32p, _c_float, _c_float, _f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# ---------------------------------------------------------------- helpers
def fft(x, inverse=False):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128))
    buf = np.ascontiguousarray(a.view(np.float64).copy())
    lib().orc_fft(buf, a.size, 1 if inverse else 0)
    return buf.view(np.complex128)


def wavelength(kv):
    return lib().orc_wavelength(kv)


def ctf_params(kv, cs_mm, ac, pix, ext_rad=0.0):
    out = np.zeros(5, np.float32)
    lib().orc_ctf_params(kv, cs_mm, ac, pix, ext_rad, out)
    return out


def rescale_sizes(nx, ny, pix):
    s = np.zeros(3, np.int32)
    p = np.zeros(2, np.float32)
    lib().orc_rescale_sizes(nx, ny, pix, s, p)
    return int(s[0]), int(s[1]), int(s[2]), float(p[0]), float(p[1])


def rescale(img, pix):
    img = np.ascontiguousarray(img, dtype=np.float32)
    ny, nx = img.shape
    nnx, nny, pitch, _, _ = rescale_sizes(nx, ny, pix)
    out = np.zeros((nny, pitch), np.float64)
    lib().orc_rescale(img, nx, ny, pix, out)
    return out


def tile_grid(nx, ny, tile=512):
    out = np.zeros(6, np.int32)
    lib().orc_tile_grid(nx, ny, tile, out)
    return tuple(int(v) for v in out)


def round_edge_table(tile):
    out = np.zeros((tile, tile), np.float64)
    lib().orc_round_edge_table(tile, out)
    return out


def avg_spectrum(pad_img, nx, ny, tile=512):
    pad_img = np.ascontiguousarray(pad_img, dtype=np.float64)
    out = np.zeros((tile, tile // 2 + 1), np.float64)
    lib().orc_avg_spectrum(pad_img, nx, ny, pad_img.shape[1], tile, out)
    return out


def rm_background(spec, log_spect=0, min_freq=1.0 / 15.0):
    spec = np.ascontiguousarray(spec, dtype=np.float64)
    out = np.zeros_like(spec)
    lib().orc_rm_background(spec, out, log_spect, spec.shape[1], spec.shape[0], min_freq)
    return out


def moments(spec):
    spec = np.ascontiguousarray(spec, dtype=np.float64)
    out = np.zeros(3, np.float64)
    lib().orc_moments(spec, spec.size, out)
    return out


def gate(spec):
    """1 when the flatness gate lets the low-pass run (CFindCtfBase.cpp:171-174)."""
    spec = np.ascontiguousarray(spec, dtype=np.float64)
    return int(lib().orc_gate(spec, spec.size))


def lowpass(spec):
    spec = np.ascontiguousarray(spec, dtype=np.float64).copy()
    lib().orc_lowpass(spec, spec.shape[1], spec.shape[0])
    return spec


def radial_table(cmpx=257, cmpy=512, use_fma=False, f64=False):
    counts = np.zeros(cmpx, np.int32)
    L = lib()
    if f64:
        n = L.orc_radial_table_f64(cmpx, cmpy, _ptr(counts), None, None)
        xy = np.zeros((n, 2), np.int32)
        dist = np.zeros(n, np.float64)
        L.orc_radial_table_f64(cmpx, cmpy, _ptr(counts), _ptr(xy), _ptr(dist))
        return counts, xy, dist
    n = L.orc_radial_table(cmpx, cmpy, int(use_fma), _ptr(counts), None)
    xy = np.zeros((n, 2), np.int32)
    L.orc_radial_table(cmpx, cmpy, int(use_fma), _ptr(counts), _ptr(xy))
    return counts, xy


def radial_avg(spec):
    spec = np.ascontiguousarray(spec, dtype=np.float64)
    out = np.zeros(spec.shape[1], np.float64)
    lib().orc_radial_avg(spec, spec.shape[1], spec.shape[0], out)
    return out


def ctf1d(wl_pix, cs_pix, df_pix, add_phase, n=257):
    out = np.zeros(n, np.float64)
    lib().orc_ctf1d(np.array([wl_pix, cs_pix], np.float32), df_pix, add_phase, n, out)
    return out


def cc1d(ctf, spec, lo, hi, bfac=1.0):
    return lib().orc_cc1d(np.ascontiguousarray(ctf, np.float64),
                          np.ascontiguousarray(spec, np.float64), len(ctf), lo, hi, bfac)


def ctf2d(wl_pix, cs_pix, dfmin_pix, dfmax_pix, az_rad, add_phase, cmpx=257, cmpy=512):
    out = np.zeros((cmpy, cmpx), np.float64)
    lib().orc_ctf2d(np.array([wl_pix, cs_pix], np.float32), dfmin_pix, dfmax_pix, az_rad,
                    add_phase, cmpx, cmpy, out)
    return out


def ring_pixels(flo, fhi, cmpx=257, cmpy=512):
    L = lib()
    n = L.orc_ring_pixels(cmpx, cmpy, flo, fhi, None)
    idx = np.zeros(n, np.int32)
    L.orc_ring_pixels(cmpx, cmpy, flo, fhi, _ptr(idx))
    return idx


def cc2d(ctf, spec, flo, fhi, bfac=16.0):
    ctf = np.ascontiguousarray(ctf, np.float64)
    spec = np.ascontiguousarray(spec, np.float64)
    return lib().orc_cc2d(ctf, spec, spec.shape[1], spec.shape[0], flo, fhi, bfac)


def spectral_cc(ctf, spec, width=10):
    ctf = np.ascontiguousarray(ctf, np.float64)
    spec = np.ascontiguousarray(spec, np.float64)
    cc = np.zeros(spec.shape[1], np.float64)
    shell = lib().orc_spectral_cc(ctf, spec, spec.shape[1], spec.shape[0], width, _ptr(cc))
    return shell, cc


class FindCtf2D:
    """The oracle's CFindCtf2D-shaped driver (CFindCTFInc.h:541-615)."""

    def __init__(self, kv=300.0, cs=2.7, ac=0.07, pix=1.2, tile=512, log_spect=0,
                 ext_phase_deg=0.0):
        self._L = lib()
        self._h = self._L.orc_fc_create()
        self.tile = tile
        self._L.orc_fc_setup(self._h, kv, cs, ac, pix, tile, log_spect, ext_phase_deg)

    def __del__(self):
        try:
            self._L.orc_fc_destroy(self._h)
        except Exception:
            pass

    def trace(self, on=True):
        self._L.orc_fc_trace_enable(self._h, int(on))

    def get_trace(self):
        n = self._L.orc_fc_trace_count(self._h)
        out = np.zeros((n, 6), np.float32)
        if n:
            self._L.orc_fc_trace_get(self._h, out)
        return out

    def gen_half_spectrum(self, pad_img, nx, ny):
        pad_img = np.ascontiguousarray(pad_img, dtype=np.float64)
        self._L.orc_fc_gen_half_spectrum(self._h, pad_img, nx, ny, pad_img.shape[1])

    def set_half_spect(self, spect):
        self._L.orc_fc_set_half_spect(self._h, np.ascontiguousarray(spect, np.float64))

    def set_raw_spect(self, raw):
        self._L.orc_fc_set_raw_spect(self._h, np.ascontiguousarray(raw, np.float64))

    def spects(self):
        c = self.tile // 2 + 1
        raw = np.zeros((self.tile, c))
        sp = np.zeros((self.tile, c))
        rad = np.zeros(c)
        self._L.orc_fc_get_spects(self._h, _ptr(raw), _ptr(sp), _ptr(rad))
        return raw, sp, rad

    def lowpassed(self):
        return bool(self._L.orc_fc_lowpassed(self._h))

    def set_phase(self, p0, rng):
        self._L.orc_fc_set_phase(self._h, p0, rng)

    def set_search1d(self, lo=0.0, hi=0.0, step=0.0):
        """Initial Do1D scan [lo, hi] A in steps of `step` A (0: reference scan)."""
        self._L.orc_fc_set_search1d(self._h, lo, hi, step)

    def do1d(self):
        self._L.orc_fc_do1d(self._h)
        return self.results()

    def refine1d(self, df, rng):
        self._L.orc_fc_refine1d(self._h, df, rng)
        return self.results()

    def do2d(self):
        self._L.orc_fc_do2d(self._h)
        return self.results()

    def refine(self, df_mean, ast_ratio, ast_angle, ext_phase):
        f = lambda v: np.array(v, np.float32)
        self._L.orc_fc_refine(self._h, f(df_mean), f(ast_ratio), f(ast_angle), f(ext_phase))
        return self.results()

    def results(self):
        out = np.zeros(6, np.float32)
        self._L.orc_fc_results(self._h, out)
        return dict(df_min=float(out[0]), df_max=float(out[1]), azimuth=float(out[2]),
                    ext_phase=float(out[3]), score=float(out[4]), ctf_res=float(out[5]))

    def full_spectrum(self):
        c = self.tile // 2 + 1
        out = np.zeros((self.tile, (c - 1) * 2))
        self._L.orc_fc_full_spectrum(self._h, out)
        return out


def estimate(img, pix, kv=300.0, cs=2.7, ac=0.07, tile=512, log_spect=0, phase=(0.0, 0.0),
             trace=False, scan1d=None):
    """Rescale + GenHalfSpectrum + Do2D for one micrograph (the reference's
    full-search package path, CFindSeriesCtfs.cpp:108-141).  scan1d:
    (lo, hi, step) A of the initial 1D scan, None for the reference's."""
    img = np.ascontiguousarray(img, np.float32)
    ny, nx = img.shape
    nnx, nny, pitch, pixn, _ = rescale_sizes(nx, ny, pix)
    pad = rescale(img, pix)
    fc = FindCtf2D(kv, cs, ac, pixn, tile, log_spect, phase[0])
    if trace:
        fc.trace(True)
    fc.gen_half_spectrum(pad, nnx, nny)
    fc.set_phase(phase[0], max(phase[1], 0.0))
    if scan1d is not None:
        fc.set_search1d(*scan1d)
    res = fc.do2d()
    return res, fc


def grid_cc(spect, grid, pix, kv=300.0, cs=2.7, ac=0.07, ext_phase_deg=0.0):
    """FP64 CC of every grid candidate against one half spectrum.  grid: dict
    with df_lo, df_hi, n_df, ast_lo, ast_hi, n_ast, az_lo, az_hi, n_az,
    res_lo, res_hi (the gctf_grid fields)."""
    spect = np.ascontiguousarray(spect, np.float64)
    gf = np.array([grid["df_lo"], grid["df_hi"], grid["ast_lo"], grid["ast_hi"], grid["az_lo"],
                   grid["az_hi"]], np.float32)
    gi = np.array([grid["n_df"], grid["n_ast"], grid["n_az"]], np.int32)
    out = np.zeros(int(gi.prod()), np.float64)
    lib().orc_grid_cc(spect, spect.shape[1], spect.shape[0], kv, cs, ac, pix, ext_phase_deg, gf, gi,
                      grid["res_lo"], grid["res_hi"], out)
    return out


# ---- proposal network (net_oracle.c) --------------------------------------
def net_param_count():
    return int(lib().orc_net_param_count())


def net_forward(params, radial):
    """FP64 forward pass: radial [n][>=256] float32 -> raw outputs [n][4]."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    r = np.ascontiguousarray(radial, dtype=np.float32)
    out = np.zeros((r.shape[0], 4), np.float64)
    lib().orc_net_forward(p, r, r.shape[1], r.shape[0], out)
    return out


def net_loss_grad(params, radial, target, grad=True):
    """FP64 mean loss and its gradient over the parameters."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    r = np.ascontiguousarray(radial, dtype=np.float32)
    t = np.ascontiguousarray(target, dtype=np.float32)
    g = np.zeros(p.size, np.float64) if grad else None
    loss = lib().orc_net_loss_grad(p, r, r.shape[1], t, r.shape[0], g.ctypes.data if grad else None)
    return (float(loss), g) if grad else float(loss)
