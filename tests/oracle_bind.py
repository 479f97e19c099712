"""ctypes binding of the oracle (oracle/_build/liboracle.so) for the tests.

The oracle is the CPU restatement of the reference used as the checker; it is
never part of the product path."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
CLI = os.path.join(ROOT, "oracle", "_build", "oracle_cli")

_L = None


class HiiParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("dx", ctypes.c_double), ("dk", ctypes.c_double),
                ("nr", ctypes.c_int), ("radii", ctypes.POINTER(ctypes.c_double)),
                ("precision", ctypes.c_int), ("dilation", ctypes.c_int)]


class Cosmo(ctypes.Structure):
    _fields_ = [("hubble", ctypes.c_double), ("omega_m", ctypes.c_double),
                ("omega_b", ctypes.c_double), ("omega_lambda", ctypes.c_double)]


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(LIB)
        d = ctypes.c_double
        for name in ("orc_Rion", "orc_Rrec", "orc_Qion", "orc_G_H", "orc_XHI", "orc_W_filter"):
            getattr(L, name).restype = d
        L.orc_Rion.argtypes = [ctypes.c_float, d]
        L.orc_Rrec.argtypes = [ctypes.c_float, d, d, ctypes.c_long, d]
        L.orc_Qion.argtypes = [d]
        L.orc_G_H.argtypes = [d]
        L.orc_XHI.argtypes = [d]
        L.orc_W_filter.argtypes = [d]
        L.orc_check_borders.restype = ctypes.c_long
        L.orc_check_borders.argtypes = [ctypes.c_long, ctypes.c_long]
        L.orc_d2max.restype = ctypes.c_long
        L.orc_d2max.argtypes = [d, d]
        L.orc_ladder.argtypes = [d, d, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_prepass_sfr.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p, d, d, d, d,
                                      ctypes.c_long, d, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_hii_slice.argtypes = [ctypes.POINTER(HiiParams), ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(d), ctypes.c_void_p]
        L.orc_filtered_d.argtypes = [ctypes.c_int, d, d, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
        L.orc_filtered_f.argtypes = L.orc_filtered_d.argtypes
        L.orc_dilate_weighted_brute.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_dilate_weighted_edt.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_dilate_weighted_maxplus.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_power_spectrum.argtypes = [ctypes.c_int, d, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_halo_emissivity.argtypes = [ctypes.c_long, ctypes.c_void_p, d, ctypes.c_long, ctypes.c_long,
                                          d, d, d, ctypes.c_void_p]
        for nm in ("orc_r2c3d_d", "orc_c2r3d_d", "orc_r2c3d_f", "orc_c2r3d_f"):
            getattr(L, nm).argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_c2c3d_d.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        for nm in ("orc_hubble_z", "orc_growth", "orc_dgrowth_dt"):
            getattr(L, nm).restype = d
            getattr(L, nm).argtypes = [ctypes.POINTER(Cosmo), d]
        L.orc_sfrd.argtypes = [ctypes.c_long, ctypes.c_void_p, d, ctypes.c_long, ctypes.c_long, d,
                               ctypes.c_void_p, ctypes.POINTER(d)]
        L.orc_nldensity.argtypes = [ctypes.c_long, ctypes.c_long, d, d, d, d] + [ctypes.c_void_p] * 6
        L.orc_t21.argtypes = [ctypes.c_int, ctypes.POINTER(Cosmo), ctypes.c_int, d] + [ctypes.c_void_p] * 6 + \
            [ctypes.POINTER(d), ctypes.POINTER(d)]
        _L = L
    return _L


def _p(a):
    return None if a is None else a.ctypes.data


def hii_slice(n, box_len, radii, rec, ion, fresid=None, precision=64, dilation=1):
    """Returns (xhii, neutral, first_centre) from the oracle."""
    r = (ctypes.c_double * len(radii))(*radii)
    p = HiiParams(n, box_len / n, 2 * 3.1415926535897932385 / box_len, len(radii), r, precision, dilation)
    rec = np.ascontiguousarray(rec, np.float32).ravel()
    ion = np.ascontiguousarray(ion, np.float32).ravel()
    fr = None if fresid is None else np.ascontiguousarray(fresid, np.float32).ravel()
    x = np.zeros(n ** 3, np.float32)
    fc = np.zeros(n ** 3, np.int32)
    neu = ctypes.c_double()
    rc = lib().orc_hii_slice(ctypes.byref(p), _p(rec), _p(ion), _p(fr), _p(x), ctypes.byref(neu), _p(fc))
    assert rc == 0
    return x, neu.value, fc


def filtered(n, box_len, R, rec, ion, precision=64):
    rec = np.ascontiguousarray(rec, np.float32).ravel()
    ion = np.ascontiguousarray(ion, np.float32).ravel()
    dt = np.float64 if precision == 64 else np.float32
    a = np.zeros(n ** 3, dt)
    b = np.zeros(n ** 3, dt)
    fn = lib().orc_filtered_d if precision == 64 else lib().orc_filtered_f
    fn(n, 2 * 3.1415926535897932385 / box_len, R, _p(rec), _p(ion), _p(a), _p(b))
    return a, b


def d2max(R, dx):
    return int(lib().orc_d2max(R, dx))


def dilate_brute(n, w):
    w = np.ascontiguousarray(w, np.int32).ravel()
    out = np.zeros(n ** 3, np.uint8)
    lib().orc_dilate_weighted_brute(n, _p(w), _p(out))
    return out


def dilate_maxplus(n, w):
    w = np.ascontiguousarray(w, np.int32).ravel()
    out = np.zeros(n ** 3, np.uint8)
    lib().orc_dilate_weighted_maxplus(n, _p(w), _p(out))
    return out


def dilate_edt(n, w):
    w = np.ascontiguousarray(w, np.int32).ravel()
    out = np.zeros(n ** 3, np.uint8)
    lib().orc_dilate_weighted_edt(n, _p(w), _p(out))
    return out


def prepass_sfr(delta, sfrd, z, fesc, box_len_h, n, hubble):
    n3 = n ** 3
    delta = np.ascontiguousarray(delta, np.float32).ravel()
    sfrd = np.ascontiguousarray(sfrd, np.float32).ravel()
    fr = np.zeros(n3, np.float32)
    rec = np.zeros(n3, np.float32)
    ion = np.zeros(n3, np.float32)
    lib().orc_prepass_sfr(n3, _p(delta), _p(sfrd), z, fesc, box_len_h / n, box_len_h, n, hubble,
                          _p(fr), _p(rec), _p(ion))
    return fr, rec, ion


def ladder(rmax, dx, nbins):
    cap = 64 * nbins + 64
    buf = np.zeros(cap)
    nm = ctypes.c_int()
    k = lib().orc_ladder(rmax, dx, nbins, _p(buf), cap, ctypes.byref(nm))
    return list(buf[:k]), nm.value


def cosmo(hubble, omega_m, omega_b, omega_lambda):
    return Cosmo(hubble, omega_m, omega_b, omega_lambda)


def growth(c, z):
    return lib().orc_growth(ctypes.byref(c), z)


def dgrowth_dt(c, z):
    return lib().orc_dgrowth_dt(ctypes.byref(c), z)


def hubble_z(c, z):
    return lib().orc_hubble_z(ctypes.byref(c), z)


def t21(n, c, vel_comp, z, delta, xhii, xc=None, xalpha=None, tk=None):
    """FP64 restatement of t21.c for one redshift: (box float32, mean, mean TS)."""
    delta = np.ascontiguousarray(delta, np.float32).ravel()
    xhii = np.ascontiguousarray(xhii, np.float32).ravel()
    sp = [None if a is None else np.ascontiguousarray(a, np.float32).ravel() for a in (xc, xalpha, tk)]
    out = np.zeros(n ** 3, np.float32)
    mt, ms = ctypes.c_double(), ctypes.c_double()
    lib().orc_t21(n, ctypes.byref(c), vel_comp, z, _p(delta), _p(xhii), _p(sp[0]), _p(sp[1]), _p(sp[2]),
                  _p(out), ctypes.byref(mt), ctypes.byref(ms))
    return out, mt.value, ms.value


def sfrd(halos, z, nh, n, dx_halo):
    halos = np.ascontiguousarray(halos, np.float32)
    out = np.zeros(n ** 3, np.float32)
    mean = ctypes.c_double()
    lib().orc_sfrd(len(halos), _p(halos), z, nh, n, dx_halo, _p(out), ctypes.byref(mean))
    return out, mean.value


def nldensity(nh, n, L_h, hubble, growth_in, growth, vx, vy, vz, delta0):
    a = [np.ascontiguousarray(x, np.float32).ravel() for x in (vx, vy, vz, delta0)]
    box = np.zeros(nh ** 3, np.float32)
    out = np.zeros(n ** 3, np.float32)
    lib().orc_nldensity(nh, n, L_h, hubble, growth_in, growth, _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]),
                        _p(box), _p(out))
    return box, out
