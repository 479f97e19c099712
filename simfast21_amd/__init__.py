"""MI355X-native ionization-field step of SimFast21 (the excursion-set HII
bubble finder of get_HIIbubbles.c).

The compute path is libsf21hii.so (hand-written HIP kernels for gfx950 behind
the C-ABI in include/sf21_hii.h); this package is a thin ctypes binding used
by the tests, the benchmark and Python callers.  There is no CPU fallback:
importing works without a GPU, but every compute call needs the built library
and a HIP device and raises otherwise.
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

import numpy as np

from . import synth  # noqa: F401

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "_build" / "libsf21hii.so"
CLI_PATH = PKG_DIR / "_build" / "get_HIIbubbles"
HEADER = PKG_DIR.parent / "include" / "sf21_hii.h"

SF21_OK = 0
SF21_ERR_ARG = -1
SF21_ERR_ALLOC = -2
SF21_ERR_HIP = -3
SF21_ERR_IO = -4
SF21_ERR_FORMAT = -5

INPUT_SFR = 0
INPUT_DIRECT = 1
INPUT_HALO = 2
PATH_FUSED = 0
PATH_PLAIN = 1
PATH_DEPOSIT = 2  # deposits only: any n, no transforms (sfrd, nldensity, halo emissivity)

FFT_R2C, FFT_C2R, FFT_C2C_FORWARD, FFT_C2C_BACKWARD = 0, 1, 2, 3


class Sf21Error(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {_lib().sf21_strerror(code).decode()} ({code})")


class HiiConfig(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("box_len", ctypes.c_double),
        ("n_radii", ctypes.c_int),
        ("radii", ctypes.POINTER(ctypes.c_double)),
        ("input_mode", ctypes.c_int),
        ("path", ctypes.c_int),
        ("hubble", ctypes.c_double),
        ("fesc", ctypes.c_double),
        ("radius_batch", ctypes.c_int),
        ("device", ctypes.c_int),
        ("filter_tile", ctypes.c_int),
        ("filter_w", ctypes.c_int),
        ("n_halo", ctypes.c_int),
    ]


class Params(ctypes.Structure):
    _fields_ = [
        ("nthreads", ctypes.c_int),
        ("n_halo", ctypes.c_long),
        ("n", ctypes.c_long),
        ("zmin", ctypes.c_double),
        ("zmax", ctypes.c_double),
        ("dz", ctypes.c_double),
        ("box_len", ctypes.c_double),
        ("hubble", ctypes.c_double),
        ("fesc", ctypes.c_double),
        ("ion_cutoff", ctypes.c_double),
        ("bubble_rmax", ctypes.c_double),
        ("bubble_nbins", ctypes.c_int),
        ("use_sfr", ctypes.c_int),
        ("dx", ctypes.c_double),
        ("dx_halo", ctypes.c_double),
        ("dk", ctypes.c_double),
        ("omega_m", ctypes.c_double),
        ("omega_b", ctypes.c_double),
        ("omega_lambda", ctypes.c_double),
        ("vel_comp", ctypes.c_int),
        ("use_lya_xrays", ctypes.c_int),
        ("zmin_sfr", ctypes.c_double),
        ("save_original_nl", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_LIB = None


def _lib():
    """Load libsf21hii.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C simfast21_amd`")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, ip, dp, i, d = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_double
    L.sf21_strerror.restype = ctypes.c_char_p
    L.sf21_strerror.argtypes = [i]
    L.sf21_hii_create.argtypes = [ctypes.POINTER(HiiConfig), ctypes.POINTER(vp)]
    L.sf21_hii_destroy.argtypes = [vp]
    L.sf21_hii_set_stream.argtypes = [vp, vp]
    L.sf21_hii_run_slice.argtypes = [vp, vp, vp, vp, d, vp, dp]
    L.sf21_hii_run_slice_dev.argtypes = [vp, vp, vp, vp, d, vp, dp]
    L.sf21_hii_run_batch_dev.argtypes = [vp, i, vp, vp, vp, dp, vp, dp]
    L.sf21_hii_weight_map.argtypes = [vp, vp]
    L.sf21_hii_dilate_dev.argtypes = [vp, vp, vp, vp, vp, d, vp, dp]
    L.sf21_hii_graph_captures.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
    L.sf21_hii_run_multi_dev.argtypes = [i, vp, vp, vp, vp, dp, vp, dp]
    L.sf21_hii_filtered_dev.argtypes = [vp, vp, vp, vp, d, i, vp]
    L.sf21_hii_profile.argtypes = [vp, i]
    L.sf21_hii_radius_batch.argtypes = [vp]
    L.sf21_hii_line_order.argtypes = [vp]
    L.sf21_hii_autotune.argtypes = [vp, vp, vp, vp, d, ctypes.c_char_p, dp]
    L.sf21_hii_tuning.argtypes = [vp, ip, ip, ip, ip, ip, ip, ip]
    L.sf21_hii_kernel_stats.argtypes = [vp, i, ctypes.c_char_p, i, ctypes.POINTER(ctypes.c_long), dp]
    L.sf21_hii_reset_stats.argtypes = [vp]
    L.sf21_hii_power_spectrum_dev.argtypes = [vp, vp, i, vp, vp, vp]
    L.sf21_hii_halo_emissivity_dev.argtypes = [vp, vp, ctypes.c_long, d, vp]
    L.sf21_hii_diff_store_fields.argtypes = [vp, i]
    L.sf21_hii_diff_last_kept.argtypes = [vp]
    L.sf21_sfrd_dev.argtypes = [vp, vp, ctypes.c_long, d, vp, dp]
    L.sf21_nldensity_dev.argtypes = [vp, ctypes.POINTER(Params), d, vp, vp, vp, vp, vp, vp]
    L.sf21_t21_dev.argtypes = [vp, ctypes.POINTER(Params), d, vp, vp, vp, vp, vp, vp, dp, dp]
    L.sf21_hii_diff_forward_dev.argtypes = [vp, vp, vp, d, i, vp, dp, vp]
    L.sf21_hii_diff_backward_dev.argtypes = [vp, vp, vp, d, d, i, d, vp, vp, dp, dp, vp]
    L.sf21_hii_run_slice_halo.argtypes = [vp, vp, vp, ctypes.c_long, d, vp, dp]
    L.sf21_fft_plan_3d.argtypes = [i, i, ctypes.POINTER(vp)]
    L.sf21_fft_execute.argtypes = [vp, vp, vp, vp]
    L.sf21_fft_destroy.argtypes = [vp]
    L.sf21_params_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(Params), ctypes.c_char_p, i]
    L.sf21_ladder.argtypes = [d, d, i, dp, i, ip]
    for fn in ("sf21_hubble_z", "sf21_growth", "sf21_dgrowth_dt"):
        getattr(L, fn).restype = d
        getattr(L, fn).argtypes = [ctypes.POINTER(Params), d]
    L.sf21_ball_d2max.restype = ctypes.c_long
    L.sf21_ball_d2max.argtypes = [d, d]
    _LIB = L
    return L


def header_symbols() -> list[str]:
    """Names of the functions declared in include/sf21_hii.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(sf21_\w+)\s*\(", text, re.M)))


def _check(rc: int, what: str):
    if rc != SF21_OK:
        raise Sf21Error(rc, what)


# ---------------------------------------------------------------------------
# host logic
# ---------------------------------------------------------------------------
def read_params(base_dir: str | os.PathLike) -> dict:
    """simfast21.ini as Input_variables.c:16-141 reads it (plus derived values)."""
    p = Params()
    msg = ctypes.create_string_buffer(512)
    rc = _lib().sf21_params_read(str(base_dir).encode(), ctypes.byref(p), msg, 512)
    if rc != SF21_OK:
        err = Sf21Error(rc, "sf21_params_read")
        err.message = msg.value.decode(errors="replace")
        raise err
    return p.as_dict()


def make_params(**kw) -> Params:
    """A Params record from keyword values (fields not given are 0)."""
    p = Params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def cosmology(z: float, hubble: float, omega_m: float, omega_lambda: float,
              omega_b: float = 0.04) -> tuple[float, float, float]:
    """(H(z) [1/s], growth D(z), dD/dt [1/s]) of auxiliary.c:224-292."""
    p = make_params(hubble=hubble, omega_m=omega_m, omega_lambda=omega_lambda, omega_b=omega_b)
    L = _lib()
    return (L.sf21_hubble_z(ctypes.byref(p), z), L.sf21_growth(ctypes.byref(p), z),
            L.sf21_dgrowth_dt(ctypes.byref(p), z))


def ladder(rmax: float, dx: float, nbins: int) -> tuple[list[float], int]:
    """Radius ladder of get_HIIbubbles.c:192-461; returns (radii, n_main)."""
    cap = 64 * max(1, nbins) + 64
    buf = (ctypes.c_double * cap)()
    nm = ctypes.c_int()
    k = _lib().sf21_ladder(rmax, dx, nbins, buf, cap, ctypes.byref(nm))
    if k < 0:
        raise Sf21Error(k, "sf21_ladder")
    return list(buf[:k]), nm.value


def ball_d2max(R: float, dx: float) -> int:
    return int(_lib().sf21_ball_d2max(R, dx))


# ---------------------------------------------------------------------------
# device helpers (torch is used for device memory only)
# ---------------------------------------------------------------------------
def _ptr(t) -> int | None:
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


class HiiContext:
    """One bubble-finder context (C-ABI sf21_hii_ctx)."""

    def __init__(self, n: int, box_len: float, radii, input_mode: int = INPUT_DIRECT,
                 path: int = PATH_FUSED, hubble: float = 0.7, fesc: float = 1.0,
                 radius_batch: int = 0, device: int = 0, filter_tile: int = 0, filter_w: int = 0,
                 n_halo: int = 0):
        self.n = int(n)
        self.radii = [float(r) for r in radii]
        self._radii = (ctypes.c_double * len(self.radii))(*self.radii)
        cfg = HiiConfig(self.n, float(box_len), len(self.radii), self._radii, input_mode, path,
                        float(hubble), float(fesc), int(radius_batch), int(device), int(filter_tile),
                        int(filter_w), int(n_halo))
        self._h = ctypes.c_void_p()
        _check(_lib().sf21_hii_create(ctypes.byref(cfg), ctypes.byref(self._h)), "sf21_hii_create")
        self.path = path
        self.input_mode = input_mode

    def close(self):
        if self._h:
            _lib().sf21_hii_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int):
        _check(_lib().sf21_hii_set_stream(self._h, stream_handle), "sf21_hii_set_stream")

    def run_slice(self, a, b, c=None, z: float = 0.0, xhii=None):
        """Device tensors (torch, float32, contiguous) -> (xhii, neutral)."""
        import torch
        if xhii is None:
            xhii = torch.empty_like(a)
        torch.cuda.current_stream().synchronize()
        neu = ctypes.c_double()
        _check(_lib().sf21_hii_run_slice_dev(self._h, _ptr(a), _ptr(b), _ptr(c), float(z), _ptr(xhii),
                                             ctypes.byref(neu)), "sf21_hii_run_slice_dev")
        return xhii, neu.value

    def run_slice_host(self, a: np.ndarray, b: np.ndarray, c: np.ndarray | None = None,
                       z: float = 0.0):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        c = None if c is None else np.ascontiguousarray(c, np.float32)
        out = np.empty_like(a)
        neu = ctypes.c_double()
        _check(_lib().sf21_hii_run_slice(self._h, _ptr(a), _ptr(b), _ptr(c), float(z), _ptr(out),
                                         ctypes.byref(neu)), "sf21_hii_run_slice")
        return out, neu.value

    def run_batch(self, a_list, b_list, c_list=None, z_list=None, out_list=None):
        import torch
        ns = len(a_list)
        if out_list is None:
            out_list = [torch.empty_like(a) for a in a_list]
        torch.cuda.current_stream().synchronize()
        P = ctypes.c_void_p * ns
        pa = P(*[_ptr(t) for t in a_list])
        pb = P(*[_ptr(t) for t in b_list])
        pc = P(*[_ptr(t) for t in c_list]) if c_list is not None else None
        po = P(*[_ptr(t) for t in out_list])
        zs = (ctypes.c_double * ns)(*([0.0] * ns if z_list is None else z_list))
        neu = (ctypes.c_double * ns)()
        _check(_lib().sf21_hii_run_batch_dev(self._h, ns, pa, pb, pc, zs, po, neu),
               "sf21_hii_run_batch_dev")
        return out_list, list(neu)

    def halo_emissivity(self, halos, z: float, out=None):
        """Device catalog tensor (n, 4) float32 -> emissivity box (device)."""
        import torch
        if out is None:
            out = torch.empty(self.n ** 3, dtype=torch.float32, device=halos.device)
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_hii_halo_emissivity_dev(self._h, _ptr(halos), int(halos.shape[0]), float(z),
                                                   _ptr(out)), "sf21_hii_halo_emissivity_dev")
        return out

    def power_spectrum(self, field, nbins: int = 20):
        """Binned P(k) of a real device box: (pk, kmean, counts) numpy arrays."""
        import torch
        pk = np.zeros(nbins)
        km = np.zeros(nbins)
        ct = np.zeros(nbins)
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_hii_power_spectrum_dev(self._h, _ptr(field), int(nbins), pk.ctypes.data,
                                                  km.ctypes.data, ct.ctypes.data),
               "sf21_hii_power_spectrum_dev")
        return pk, km, ct

    def diff_store_fields(self, mode: int = 1):
        """Backward field storage: 1 keep the first sweep's per-radius pairs
        (n_R n^3 complex floats) for the second, 0 recompute them, 2 keep them
        if they fit in half the free device memory (the default), 3 the exact
        forward keeps them for the next backward on the same inputs."""
        _check(_lib().sf21_hii_diff_store_fields(self._h, int(mode)), "sf21_hii_diff_store_fields")

    @property
    def diff_last_kept(self) -> bool:
        """Whether the last diff_backward read fields kept by the forward."""
        return bool(_lib().sf21_hii_diff_last_kept(self._h))

    def diff_forward(self, g, emis, zeta: float, nbins: int = 20, x_out=None):
        """Exact forward of the differentiable model on device boxes g, emis:
        (x, mean_x, pk).  See sf21_hii_diff_forward_dev."""
        import torch
        if x_out is None:
            x_out = torch.empty_like(g)
        pk = np.zeros(nbins)
        mx = ctypes.c_double()
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_hii_diff_forward_dev(self._h, _ptr(g), _ptr(emis), float(zeta), int(nbins),
                                                _ptr(x_out), ctypes.byref(mx), pk.ctypes.data),
               "sf21_hii_diff_forward_dev")
        return x_out, mx.value, pk

    def diff_backward(self, g, emis, zeta: float, tau: float, dl_dx: float, dl_dpk,
                      nbins: int = 20, grad_out=None):
        """Relaxed-model gradient of dl_dx <x~> + sum dl_dpk P_b:
        (grad_g (device), grad_zeta, relaxed mean x~, relaxed pk)."""
        import torch
        if grad_out is None:
            grad_out = torch.empty_like(g)
        gp = np.ascontiguousarray(np.zeros(nbins) if dl_dpk is None else dl_dpk, np.float64)
        pk = np.zeros(nbins)
        gz = ctypes.c_double()
        mx = ctypes.c_double()
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_hii_diff_backward_dev(self._h, _ptr(g), _ptr(emis), float(zeta), float(tau),
                                                 int(nbins), float(dl_dx), gp.ctypes.data,
                                                 _ptr(grad_out), ctypes.byref(gz), ctypes.byref(mx),
                                                 pk.ctypes.data),
               "sf21_hii_diff_backward_dev")
        return grad_out, gz.value, mx.value, pk

    def sfrd(self, halos, z: float, out=None):
        """SFRD box of one redshift from a device halo catalog (n, 4):
        (box, mean).  The context needs n_halo."""
        import torch
        if out is None:
            out = torch.empty(self.n ** 3, dtype=torch.float32, device=halos.device)
        mean = ctypes.c_double()
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_sfrd_dev(self._h, _ptr(halos), int(halos.shape[0]), float(z), _ptr(out),
                                    ctypes.byref(mean)), "sf21_sfrd_dev")
        return out, mean.value

    def nldensity(self, params, z: float, vx, vy, vz, delta0, box_halo=None, out=None):
        """Non-linear density of one redshift (sf21_nldensity_dev): returns
        (out (n^3), box_halo (N_halo^3) or None)."""
        import torch
        p = params if isinstance(params, Params) else make_params(**params)
        if out is None:
            out = torch.empty(self.n ** 3, dtype=torch.float32, device=vx.device)
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_nldensity_dev(self._h, ctypes.byref(p), float(z), _ptr(vx), _ptr(vy), _ptr(vz),
                                         _ptr(delta0), _ptr(box_halo), _ptr(out)), "sf21_nldensity_dev")
        return out, box_halo

    def t21(self, params, z: float, delta, xhii, xc=None, xalpha=None, tk=None, out=None):
        """21-cm brightness temperature of one redshift (sf21_t21_dev) on device
        boxes; params: a Params or a dict of its fields.  Returns (box, mean
        t21, mean TS)."""
        import torch
        p = params if isinstance(params, Params) else make_params(**params)
        if out is None:
            out = torch.empty_like(delta)
        mt, ms = ctypes.c_double(), ctypes.c_double()
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_t21_dev(self._h, ctypes.byref(p), float(z), _ptr(delta), _ptr(xhii), _ptr(xc),
                                   _ptr(xalpha), _ptr(tk), _ptr(out), ctypes.byref(mt), ctypes.byref(ms)),
               "sf21_t21_dev")
        return out, mt.value, ms.value

    def dilate(self, wmap: np.ndarray, a, b, c=None, z: float = 0.0, xhii=None):
        """The dilation and final pass on a given per-cell ball-bound map
        (host, n^3 ints, -1 = no centre); device inputs as run_slice ->
        (xhii, neutral)."""
        import torch
        w = np.ascontiguousarray(wmap, np.int32).ravel()
        if w.size != self.n ** 3:
            raise ValueError("map size")
        if xhii is None:
            xhii = torch.empty_like(a)
        torch.cuda.current_stream().synchronize()
        neu = ctypes.c_double()
        _check(_lib().sf21_hii_dilate_dev(self._h, w.ctypes.data, _ptr(a), _ptr(b), _ptr(c), float(z),
                                          _ptr(xhii), ctypes.byref(neu)), "sf21_hii_dilate_dev")
        return xhii, neu.value

    def weight_map(self) -> np.ndarray:
        out = np.empty(self.n ** 3, np.int32)
        _check(_lib().sf21_hii_weight_map(self._h, out.ctypes.data), "sf21_hii_weight_map")
        return out

    @property
    def graph_captures(self) -> int:
        """hipGraphs captured so far with this context as their owner."""
        v = ctypes.c_long()
        _check(_lib().sf21_hii_graph_captures(self._h, ctypes.byref(v)), "sf21_hii_graph_captures")
        return v.value

    def filtered(self, a, b, c=None, z: float = 0.0, radius_index: int = 0):
        import torch
        out = torch.empty(2 * self.n ** 3, dtype=torch.float32, device=a.device)
        torch.cuda.current_stream().synchronize()
        _check(_lib().sf21_hii_filtered_dev(self._h, _ptr(a), _ptr(b), _ptr(c), float(z),
                                            int(radius_index), _ptr(out)), "sf21_hii_filtered_dev")
        return out.view(-1, 2)

    def autotune(self, a, b, c=None, z: float = 0.0, cache_path: str | None = None) -> float:
        """Time the fused path's candidates on this slice, keep the fastest;
        returns its ms per slice."""
        import torch
        torch.cuda.current_stream().synchronize()
        ms = ctypes.c_double()
        _check(_lib().sf21_hii_autotune(self._h, _ptr(a), _ptr(b), _ptr(c), float(z),
                                        None if cache_path is None else str(cache_path).encode(),
                                        ctypes.byref(ms)), "sf21_hii_autotune")
        return ms.value

    def tuning(self) -> dict:
        v = [ctypes.c_int() for _ in range(7)]
        _check(_lib().sf21_hii_tuning(self._h, *[ctypes.byref(x) for x in v]), "sf21_hii_tuning")
        return {"filter_tile": v[0].value, "radius_batch": v[1].value, "filter_w": v[2].value,
                "cols_tile": v[3].value, "cols_wide": v[4].value, "rows_wide": v[5].value,
                "filter_pair": v[6].value}

    @property
    def radius_batch(self) -> int:
        return int(_lib().sf21_hii_radius_batch(self._h))

    @property
    def line_order(self) -> int:
        """Sides without a register plan: the XCD-order mask of the line
        passes (bit 1: along j, bit 2: along i)."""
        v = int(_lib().sf21_hii_line_order(self._h))
        _check(min(v, 0), "sf21_hii_line_order")
        return v

    def profile(self, enable: bool = True):
        _check(_lib().sf21_hii_profile(self._h, int(enable)), "sf21_hii_profile")

    def reset_stats(self):
        _check(_lib().sf21_hii_reset_stats(self._h), "sf21_hii_reset_stats")

    def kernel_stats(self) -> dict:
        L = _lib()
        n = L.sf21_hii_kernel_stats(self._h, -1, None, 0, None, None)
        out = {}
        name = ctypes.create_string_buffer(64)
        cnt = ctypes.c_long()
        ms = ctypes.c_double()
        for k in range(n):
            _check(L.sf21_hii_kernel_stats(self._h, k, name, 64, ctypes.byref(cnt), ctypes.byref(ms)),
                   "sf21_hii_kernel_stats")
            out[name.value.decode()] = (cnt.value, ms.value)
        return out


def run_multi(contexts, a_list, b_list, c_list=None, z_list=None, out_list=None):
    """Boxes of different sizes in one hipGraph launch (sf21_hii_run_multi_dev):
    box i runs on contexts[i] (a HiiContext of its size).  Returns (outs,
    neutral list)."""
    import torch
    nb = len(a_list)
    if out_list is None:
        out_list = [torch.empty_like(a) for a in a_list]
    torch.cuda.current_stream().synchronize()
    P = ctypes.c_void_p * nb
    pc = P(*[c._h.value for c in contexts])
    pa = P(*[_ptr(t) for t in a_list])
    pb = P(*[_ptr(t) for t in b_list])
    pcc = P(*[_ptr(t) for t in c_list]) if c_list is not None else None
    po = P(*[_ptr(t) for t in out_list])
    zs = (ctypes.c_double * nb)(*([0.0] * nb if z_list is None else z_list))
    neu = (ctypes.c_double * nb)()
    _check(_lib().sf21_hii_run_multi_dev(nb, pc, pa, pb, pcc, zs, po, neu), "sf21_hii_run_multi_dev")
    return out_list, list(neu)


class FftPlan:
    """FFTW-like 3D transform on device memory (C-ABI sf21_fft_*)."""

    def __init__(self, n: int, kind: int):
        self.n, self.kind = n, kind
        self._h = ctypes.c_void_p()
        _check(_lib().sf21_fft_plan_3d(n, kind, ctypes.byref(self._h)), "sf21_fft_plan_3d")

    def execute(self, inp, out, stream: int = 0):
        _check(_lib().sf21_fft_execute(self._h, _ptr(inp), _ptr(out), stream or None), "sf21_fft_execute")

    def __del__(self):
        try:
            if self._h:
                _lib().sf21_fft_destroy(self._h)
        except Exception:
            pass
