"""ctypes binding of ``lib/libpdspy_amd.so`` (the C ABI in include/pdspy_amd.h).

There is no fallback: if the library or a GPU is missing, every entry point
raises.  Device memory and streams come from PyTorch (plumbing only).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PDSPY_AMD_LIB selects another build of the same ABI (e.g. a profiling build)
LIB_PATH = os.environ.get("PDSPY_AMD_LIB") or os.path.join(HERE, "lib", "libpdspy_amd.so")

# symbol list (must match include/pdspy_amd.h)
EXPORTS = [
    "rt_create", "rt_destroy", "rt_set_stream", "rt_last_error", "rt_device_info",
    "rt_model_create", "rt_model_destroy", "rt_model_ncells", "rt_model_set_density",
    "rt_model_set_opacity", "rt_model_set_opacity_device", "rt_model_get_tables",
    "rt_model_ngroups", "rt_model_get_opacity",
    "rt_model_generate_density", "rt_get_density_device", "rt_scattering_source",
    "rt_thermal", "rt_get_temperature", "rt_thermal_timeline", "rt_get_temperature_device",
    "rt_image",
    "rt_image_adjoint", "rt_vis_dft", "rt_vis_dft_adjoint", "rt_vis_grid",
    "rt_vis_grid_workspace", "rt_vis_grid_noalloc", "rt_interpolate_galario", "rt_chisq",
    "rt_batch_create", "rt_batch_destroy", "rt_batch_thermal", "rt_batch_step",
    "rt_refine_stats", "rt_refine_reserve", "rt_image_crossings",
]

RT_ERRORS = {-1: "RT_E_ARG", -2: "RT_E_HIP", -3: "RT_E_NOMEM", -4: "RT_E_WATCHDOG",
             -5: "RT_E_OVERFLOW", -6: "RT_E_LIMIT"}


class RTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (RT_ERRORS.get(code, "RT_E?"), code, msg))
        self.code = code


class ModelDesc(C.Structure):
    _fields_ = [("nr", C.c_int), ("ntheta", C.c_int), ("nphi", C.c_int),
                ("h_r_walls", C.c_void_p), ("h_theta_walls", C.c_void_p),
                ("h_phi_walls", C.c_void_p), ("nspec", C.c_int), ("h_rho", C.c_void_p),
                ("nlam", C.c_int), ("h_lam_um", C.c_void_p), ("h_kabs", C.c_void_p),
                ("h_ksca", C.c_void_p), ("h_g", C.c_void_p), ("nstar", C.c_int),
                ("h_stars", C.c_void_p), ("ntemp", C.c_int), ("temp0", C.c_double),
                ("temp1", C.c_double)]


RT_COMP_DISK, RT_COMP_ULRICH_EXT = 1, 2


class DensityComp(C.Structure):
    _fields_ = [("kind", C.c_int), ("species", C.c_int), ("mass", C.c_double),
                ("rmin", C.c_double), ("rmax", C.c_double), ("plrho", C.c_double),
                ("h0", C.c_double), ("plh", C.c_double), ("rcent", C.c_double),
                ("cavpl", C.c_double), ("cavrfact", C.c_double), ("theta_open", C.c_double),
                ("zoffset", C.c_double)]


class ThermalOpts(C.Structure):
    _fields_ = [("max_events", C.c_uint64), ("epoch_first", C.c_uint64),
                ("epoch_growth", C.c_double), ("count_visits", C.c_int), ("mrw", C.c_int),
                ("mrw_tauthres", C.c_double), ("mrw_count_trigger", C.c_int),
                ("mrw_gamma", C.c_double), ("max_seconds", C.c_double),
                ("block_size", C.c_int), ("blocks_per_cu", C.c_int),
                ("slice_events", C.c_uint64), ("no_park", C.c_int),
                ("free_cus_per_xcd", C.c_int)]


class ThermalStats(C.Structure):
    _fields_ = [("n_cross", C.c_uint64), ("n_abs", C.c_uint64), ("n_scat", C.c_uint64),
                ("n_escape", C.c_uint64), ("n_killed", C.c_uint64), ("n_events", C.c_uint64),
                ("n_mrw", C.c_uint64), ("eps", C.c_double), ("lum", C.c_double),
                ("n_epochs", C.c_int), ("watchdog_fired", C.c_int), ("kernel_ms", C.c_double),
                ("total_ms", C.c_double), ("max_epoch_ms", C.c_double),
                ("max_packet_steps", C.c_uint64), ("wave_iters", C.c_uint64),
                ("idle_iters", C.c_uint64), ("n_nudged", C.c_uint64),
                ("phase_clocks", C.c_uint64 * 8), ("n_parked", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["phase_clocks"] = list(self.phase_clocks)
        return d


class Camera(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nsub", C.c_int), ("pix_cm", C.c_double),
                ("incl_deg", C.c_double), ("posang_deg", C.c_double), ("phi_deg", C.c_double),
                ("lam_um", C.c_double), ("mode", C.c_int), ("incl_star", C.c_int),
                ("tau_stop", C.c_double), ("tile_w", C.c_int), ("tile_h", C.c_int),
                ("d_jscat", C.c_void_p), ("nrefine", C.c_int), ("refine_crit", C.c_double)]


class ScatStats(C.Structure):
    _fields_ = [("n_escape", C.c_uint64), ("n_abs", C.c_uint64), ("n_scat", C.c_uint64),
                ("n_cross", C.c_uint64), ("n_killed", C.c_uint64), ("lum_star", C.c_double),
                ("lum_dust", C.c_double), ("eps", C.c_double), ("kernel_ms", C.c_double)]


class StepItem(C.Structure):
    _fields_ = [("cam", Camera), ("dxy", C.c_double), ("scale", C.c_double),
                ("d_u", C.c_void_p), ("d_v", C.c_void_p), ("K", C.c_int),
                ("dra", C.c_double), ("ddec", C.c_double),
                ("d_re", C.c_void_p), ("d_im", C.c_void_p), ("d_w", C.c_void_p),
                ("d_img", C.c_void_p), ("d_chi2", C.c_void_p),
                ("d_g_logrho", C.c_void_p), ("d_g_T", C.c_void_p), ("nphot_scat", C.c_uint64)]


_lib = None


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("pdspy_amd: %s not built; run `python -m pdspy_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P, I, D, U64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64
    sig = {
        "rt_create": ([I, C.POINTER(P)], I),
        "rt_destroy": ([P], I),
        "rt_set_stream": ([P, P], I),
        "rt_last_error": ([P], C.c_char_p),
        "rt_device_info": ([P, C.POINTER(I), C.POINTER(I)], I),
        "rt_model_create": ([P, C.POINTER(ModelDesc), C.POINTER(P)], I),
        "rt_model_destroy": ([P], I),
        "rt_model_ncells": ([P], I),
        "rt_model_set_density": ([P, P], I),
        "rt_model_set_opacity": ([P, P, P, P], I),
        "rt_model_set_opacity_device": ([P, P, P, P], I),
        "rt_model_get_tables": ([P, P, P, P, P, P], I),
        "rt_model_ngroups": ([P], I),
        "rt_model_get_opacity": ([P, P, P, P, P, P, P], I),
        "rt_model_generate_density": ([P, C.POINTER(DensityComp), I], I),
        "rt_get_density_device": ([P, P], I),
        "rt_scattering_source": ([P, D, U64, U64, P, P, C.POINTER(ScatStats)], I),
        "rt_thermal": ([P, U64, U64, C.POINTER(ThermalOpts), C.POINTER(ThermalStats)], I),
        "rt_get_temperature": ([P, P, P], I),
        "rt_thermal_timeline": ([P, P, I], I),
        "rt_get_temperature_device": ([P, P], I),
        "rt_image": ([P, C.POINTER(Camera), P, P, P], I),
        "rt_image_adjoint": ([P, C.POINTER(Camera), P, P, P, P, P], I),
        "rt_vis_dft": ([P, P, I, I, D, P, P, I, D, D, P, P], I),
        "rt_vis_dft_adjoint": ([P, P, P, I, I, D, P, P, I, D, D, P], I),
        "rt_vis_grid": ([P, P, I, I, D, P, P, I, D, D, P, P, P], I),
        "rt_vis_grid_workspace": ([I, I], C.c_size_t),
        "rt_vis_grid_noalloc": ([P, P, I, I, D, P, P, I, D, D, P, P], I),
        "rt_refine_stats": ([P, C.POINTER(U64), C.POINTER(U64)], I),
        "rt_image_crossings": ([P, C.POINTER(U64), C.POINTER(U64), I], I),
        "rt_refine_reserve": ([P, U64], I),
        "rt_interpolate_galario": ([P, P, I, I, D, P, P, I, D, D, P, P, P], I),
        "rt_chisq": ([P, P, P, P, P, P, I, P, P, P], I),
        "rt_batch_create": ([P, P, I, C.POINTER(P)], I),
        "rt_batch_destroy": ([P], I),
        "rt_batch_thermal": ([P, P, P, C.POINTER(ThermalOpts), P], I),
        "rt_batch_step": ([P, P, P, C.POINTER(ThermalOpts), P, I, P, C.POINTER(C.c_double)], I),
    }
    for name in EXPORTS:
        f = getattr(L, name)
        f.argtypes, f.restype = sig[name]
    _lib = L
    return L


class Context:
    """One device + stream.  Wraps ``rt_ctx``."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("pdspy_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.L = load()
        self.device = device
        torch.cuda.set_device(device)
        h = C.c_void_p()
        rc = self.L.rt_create(device, C.byref(h))
        if rc != 0:
            raise RTError(rc, "rt_create(%d) failed" % device)
        self.h = h

    def sync_stream(self):
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        self.L.rt_set_stream(self.h, C.c_void_p(s))

    def check(self, rc):
        if rc != 0:
            msg = self.L.rt_last_error(self.h)
            raise RTError(rc, msg.decode() if msg else "")
        return rc

    def info(self):
        n, k = C.c_int(), C.c_int()
        self.L.rt_device_info(self.h, C.byref(n), C.byref(k))
        return n.value, k.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.rt_destroy(self.h)
                self.h = None
        except Exception:
            pass


_ctx = {}


def context(device=0):
    if device not in _ctx:
        _ctx[device] = Context(device)
    c = _ctx[device]
    c.sync_stream()
    return c
