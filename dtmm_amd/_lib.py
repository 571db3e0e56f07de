"""ctypes binding of the C ABI declared in ``include/dtmmx.h``.

The library is ``dtmm_amd/libdtmmx.so`` built by ``dtmm_amd.build``.  There
is no fallback: if the library is missing every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# DTMMX_LIB selects another build of the same library (e.g. a variant built
# with different compile-time options).
LIB_PATH = os.environ.get("DTMMX_LIB") or os.path.join(HERE, "libdtmmx.so")

DTMMX_OK = 0
DTMMX_E_ARG = -1
DTMMX_E_UNSUPPORTED = -2
DTMMX_E_HIP = -3
DTMMX_E_NOMEM = -4
DTMMX_C64 = 0
DTMMX_C128 = 1
DTMMX_METHOD_2X2 = 2
DTMMX_METHOD_4X4 = 4
ABI_VERSION = 5
MAX_WAVELENGTHS = 64
MAX_BEAMS = 64

# Every symbol include/dtmmx.h declares.
EXPORTS = (
    "dtmmx_abi_version",
    "dtmmx_last_error",
    "dtmmx_plan_create",
    "dtmmx_plan_destroy",
    "dtmmx_plan_workspace_size",
    "dtmmx_transfer_workspace_size",
    "dtmmx_plan_set_tiles",
    "dtmmx_plan_get_tiles",
    "dtmmx_plan_set_row_split",
    "dtmmx_plan_get_row_split",
    "dtmmx_transfer",
    "dtmmx_transfer_bulk",
    "dtmmx_fft2",
    "dtmmx_host_alphaf",
    "dtmmx_timing_begin",
    "dtmmx_timing_end",
    "dtmmx_host_rop",
    "dtmmx_host_windows",
    "dtmmx_transfer_batch",
    "dtmmx_tape_size",
    "dtmmx_adjoint_workspace_size",
    "dtmmx_transfer_record",
    "dtmmx_transfer_adjoint",
    "dtmmx_host_rop_grad",
    "dtmmx_pom_workspace_size",
    "dtmmx_pom_specter",
    "dtmmx_specter2color_workspace_size",
    "dtmmx_specter2color",
    "dtmmx_tmm_transfer4x4",
    "dtmmx_tmm_transfer",
    "dtmmx_illumination_workspace_size",
    "dtmmx_illumination",
    "dtmmx_tmm2d_workspace_size",
    "dtmmx_tmm2d_stack",
    "dtmmx_tmm2d_system",
    "dtmmx_tmm2d_reflection_workspace_size",
    "dtmmx_tmm2d_reflection",
    "dtmmx_tmm2d_reflect",
    "dtmmx_tmm2d_modes",
    "dtmmx_director2data_blocks",
    "dtmmx_director2data",
    "dtmmx_q2data",
    "dtmmx_transfer_noalloc",
    "dtmmx_host_sincos",
    "dtmmx_propagate_layer",
)

_dp = ctypes.POINTER(ctypes.c_double)


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int),
        ("width", ctypes.c_int),
        ("n_wavelengths", ctypes.c_int),
        ("n_rays", ctypes.c_int),
        ("rays_per_beam", ctypes.c_int),
        ("method", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("betamax", ctypes.c_double),
        ("k0", _dp),
        ("beta", _dp),
        ("phi", _dp),
        ("diffraction", ctypes.c_int),
        ("npass", ctypes.c_int),
        ("norm", ctypes.c_int),
        ("smooth", ctypes.c_double),
        ("reflection", ctypes.c_int),
        ("checkpoint", ctypes.c_int),
    ]


class Stack(ctypes.Structure):
    _fields_ = [
        ("n_layers", ctypes.c_int),
        ("thickness", _dp),
        ("eps_eff_re", _dp),
        ("eps_eff_im", _dp),
        ("eps_re", ctypes.c_void_p),
        ("eps_im", ctypes.c_void_p),
        ("angles", ctypes.c_void_p),
        ("n_in", ctypes.c_double),
        ("n_out", ctypes.c_double),
        ("n_steps", ctypes.POINTER(ctypes.c_int)),
        ("biaxial", ctypes.POINTER(ctypes.c_int)),
        ("eff_eps_re", _dp),
        ("eff_eps_im", _dp),
        ("eff_angles", _dp),
        ("eff_thickness", _dp),
    ]


class Job(ctypes.Structure):
    _fields_ = [
        ("plan", ctypes.c_void_p),
        ("stack", ctypes.POINTER(Stack)),
        ("field_in", ctypes.c_void_p),
        ("field_out", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p),
        ("workspace_bytes", ctypes.c_size_t),
    ]


class PomDesc(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int),
        ("width", ctypes.c_int),
        ("n_wavelengths", ctypes.c_int),
        ("k0", _dp),
        ("n_fields", ctypes.c_int),
        ("n_pol", ctypes.c_int),
        ("jones", ctypes.c_double * 4),
        ("diffraction", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("d", ctypes.c_double),
        ("eps_re", ctypes.c_double),
        ("eps_im", ctypes.c_double),
        ("betamax", ctypes.c_double),
        ("pmat", _dp),
        ("n_pmat", ctypes.c_int),
        ("precision", ctypes.c_int),
    ]


class TmmDesc(ctypes.Structure):
    _fields_ = [
        ("n_elements", ctypes.c_longlong),
        ("n_layers", ctypes.c_int),
        ("kd", ctypes.c_void_p),
        ("eps_re", ctypes.c_void_p),
        ("eps_im", ctypes.c_void_p),
        ("angles", ctypes.c_void_p),
        ("beta", ctypes.c_double),
        ("phi", ctypes.c_double),
        ("n_in", ctypes.c_double),
        ("n_out", ctypes.c_double),
        ("method", ctypes.c_int),
        ("cfact", ctypes.c_double),
        ("stack_beta", ctypes.c_double),
        ("stack_phi", ctypes.c_double),
        ("reflect_in", ctypes.c_int),
        ("reflect_out", ctypes.c_int),
    ]


TMM_METHODS = {"4x4": 0, "4x4_1": 1, "4x4_2": 2, "4x4_r": 3, "2x2": 4, "2x2_1": 5}


class Tmm2dDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int),
        ("n_layers", ctypes.c_int),
        ("n_modes", ctypes.c_int),
        ("mode_index", ctypes.c_void_p),
        ("k0", ctypes.c_double),
        ("betay", ctypes.c_double),
        ("thickness", ctypes.c_void_p),
        ("eps_re", ctypes.c_void_p),
        ("eps_im", ctypes.c_void_p),
        ("angles", ctypes.c_void_p),
        ("method", ctypes.c_int),
    ]


class IllumDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_double),
        ("focus", ctypes.c_double),
        ("mode", ctypes.c_int),
        ("diffraction", ctypes.c_int),
        ("polarized", ctypes.c_int),
        ("jones_re", ctypes.c_double * 2),
        ("jones_im", ctypes.c_double * 2),
        ("intensity", ctypes.c_void_p),
        ("window", ctypes.c_void_p),
    ]


class DtmmxError(RuntimeError):
    """Error code and message returned by the engine library."""

    def __init__(self, code, msg):
        super().__init__("dtmmx error %d: %s" % (code, msg))
        self.code = code


_lib = None


def source_hash():
    """SHA-256 of the library's sources (dtmm_amd/csrc/*.hip, *.hpp and
    include/*.h, in name order); None when they are not present."""
    import glob
    import hashlib
    root = os.path.dirname(HERE)
    files = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")) +
                   glob.glob(os.path.join(HERE, "csrc", "*.hpp")) +
                   glob.glob(os.path.join(root, "include", "*.h")))
    if not files:
        return None
    h = hashlib.sha256()
    for f in files:
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def _check_fresh(path):
    """The build writes the hash of the sources it compiled next to the
    library; a library whose sources have changed since is refused."""
    want = source_hash()
    if want is None:
        return
    rec = path + ".srchash"
    if not os.path.exists(rec):
        raise RuntimeError("%s has no build record (%s): rebuild it with `python -m "
                           "dtmm_amd.build`" % (path, os.path.basename(rec)))
    with open(rec) as fh:
        parts = fh.read().split()
    got = parts[0] if parts else ""
    if got != want:
        raise RuntimeError("%s was built from other sources than dtmm_amd/csrc and include/ "
                           "hold now: rebuild it with `python -m dtmm_amd.build`" % path)


def load():
    """Load (once) and return the engine library; raises if it is absent or
    was built from other sources than the ones present."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "dtmm_amd engine library %s is missing: build it with "
            "`python -m dtmm_amd.build` (there is no CPU fallback)" % LIB_PATH)
    _check_fresh(LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.dtmmx_abi_version.restype = ctypes.c_int
    lib.dtmmx_last_error.restype = ctypes.c_char_p
    lib.dtmmx_plan_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(PlanDesc)]
    lib.dtmmx_plan_create.restype = ctypes.c_int
    lib.dtmmx_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.dtmmx_plan_destroy.restype = ctypes.c_int
    lib.dtmmx_plan_workspace_size.argtypes = [ctypes.c_void_p]
    lib.dtmmx_plan_workspace_size.restype = ctypes.c_size_t
    lib.dtmmx_transfer_workspace_size.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.dtmmx_transfer_workspace_size.restype = ctypes.c_size_t
    lib.dtmmx_plan_set_tiles.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.dtmmx_plan_get_tiles.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int)]
    lib.dtmmx_plan_set_row_split.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.dtmmx_plan_get_row_split.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.dtmmx_transfer.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stack), ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_void_p, ctypes.c_void_p]
    lib.dtmmx_transfer_bulk.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stack), ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    lib.dtmmx_transfer.restype = ctypes.c_int
    lib.dtmmx_transfer_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stack),
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]
    lib.dtmmx_transfer_noalloc.restype = ctypes.c_int
    lib.dtmmx_propagate_layer.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stack),
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_void_p, ctypes.c_void_p]
    lib.dtmmx_propagate_layer.restype = ctypes.c_int
    lib.dtmmx_transfer_batch.argtypes = [ctypes.POINTER(Job), ctypes.c_int, ctypes.c_void_p,
                                         ctypes.c_void_p]
    lib.dtmmx_transfer_batch.restype = ctypes.c_int
    lib.dtmmx_tape_size.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stack)]
    lib.dtmmx_tape_size.restype = ctypes.c_size_t
    lib.dtmmx_adjoint_workspace_size.argtypes = [ctypes.c_void_p]
    lib.dtmmx_adjoint_workspace_size.restype = ctypes.c_size_t
    lib.dtmmx_transfer_record.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stack), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_void_p, ctypes.c_void_p]
    lib.dtmmx_transfer_record.restype = ctypes.c_int
    lib.dtmmx_transfer_adjoint.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stack), ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    lib.dtmmx_transfer_adjoint.restype = ctypes.c_int
    lib.dtmmx_host_rop_grad.argtypes = [ctypes.c_int] + [_dp] * 9
    lib.dtmmx_host_rop_grad.restype = ctypes.c_int
    lib.dtmmx_fft2.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    lib.dtmmx_fft2.restype = ctypes.c_int
    lib.dtmmx_host_alphaf.argtypes = [ctypes.c_int] + [_dp] * 7
    lib.dtmmx_host_alphaf.restype = ctypes.c_int
    lib.dtmmx_host_rop.argtypes = [ctypes.c_int] + [_dp] * 7 + [ctypes.POINTER(ctypes.c_int)]
    lib.dtmmx_host_rop.restype = ctypes.c_int
    _fp = ctypes.POINTER(ctypes.c_float)
    lib.dtmmx_host_sincos.argtypes = [ctypes.c_int, _fp, _fp, _fp]
    lib.dtmmx_host_sincos.restype = ctypes.c_int
    _ip = ctypes.POINTER(ctypes.c_int)
    lib.dtmmx_host_windows.argtypes = [ctypes.POINTER(PlanDesc), ctypes.c_int, _ip, _ip, _ip,
                                       _dp, _dp]
    lib.dtmmx_host_windows.restype = ctypes.c_int
    lib.dtmmx_timing_begin.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.dtmmx_timing_begin.restype = ctypes.c_int
    lib.dtmmx_timing_end.argtypes = [_dp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    lib.dtmmx_timing_end.restype = ctypes.c_int
    lib.dtmmx_pom_workspace_size.argtypes = [ctypes.POINTER(PomDesc)]
    lib.dtmmx_pom_workspace_size.restype = ctypes.c_size_t
    lib.dtmmx_pom_specter.argtypes = [ctypes.POINTER(PomDesc), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.dtmmx_pom_specter.restype = ctypes.c_int
    lib.dtmmx_specter2color_workspace_size.argtypes = [ctypes.c_longlong, ctypes.c_int]
    lib.dtmmx_specter2color_workspace_size.restype = ctypes.c_size_t
    lib.dtmmx_specter2color.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                        ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.dtmmx_specter2color.restype = ctypes.c_int
    lib.dtmmx_tmm_transfer4x4.argtypes = [ctypes.POINTER(TmmDesc), ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p]
    lib.dtmmx_tmm_transfer4x4.restype = ctypes.c_int
    lib.dtmmx_tmm_transfer.argtypes = [ctypes.POINTER(TmmDesc), ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p]
    lib.dtmmx_tmm_transfer.restype = ctypes.c_int
    lib.dtmmx_illumination_workspace_size.argtypes = [ctypes.c_void_p]
    lib.dtmmx_illumination_workspace_size.restype = ctypes.c_size_t
    lib.dtmmx_illumination.argtypes = [ctypes.c_void_p, ctypes.POINTER(IllumDesc), ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.dtmmx_illumination.restype = ctypes.c_int
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.dtmmx_tmm2d_workspace_size.argtypes = [ctypes.POINTER(Tmm2dDesc)]
    lib.dtmmx_tmm2d_workspace_size.restype = sz
    lib.dtmmx_tmm2d_stack.argtypes = [ctypes.POINTER(Tmm2dDesc), vp, vp, sz, vp]
    lib.dtmmx_tmm2d_stack.restype = ctypes.c_int
    lib.dtmmx_tmm2d_system.argtypes = [ctypes.c_int, vp, vp, vp, vp, vp, sz, vp]
    lib.dtmmx_tmm2d_system.restype = ctypes.c_int
    lib.dtmmx_tmm2d_reflection_workspace_size.argtypes = [ctypes.c_int]
    lib.dtmmx_tmm2d_reflection_workspace_size.restype = sz
    lib.dtmmx_tmm2d_reflection.argtypes = [ctypes.c_int, vp, vp, vp, sz, vp]
    lib.dtmmx_tmm2d_reflection.restype = ctypes.c_int
    lib.dtmmx_tmm2d_reflect.argtypes = [ctypes.c_int, ctypes.c_longlong, vp, vp, vp, vp, vp,
                                        ctypes.c_int, vp, sz, vp]
    lib.dtmmx_tmm2d_reflect.restype = ctypes.c_int
    lib.dtmmx_tmm2d_modes.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_longlong, vp, vp,
                                      ctypes.c_int, vp, sz, vp]
    lib.dtmmx_tmm2d_modes.restype = ctypes.c_int
    lib.dtmmx_director2data_blocks.argtypes = [ctypes.c_longlong]
    lib.dtmmx_director2data_blocks.restype = ctypes.c_longlong
    lib.dtmmx_director2data.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.dtmmx_director2data.restype = ctypes.c_int
    lib.dtmmx_q2data.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p]
    lib.dtmmx_q2data.restype = ctypes.c_int
    if lib.dtmmx_abi_version() != ABI_VERSION:
        raise RuntimeError("dtmmx ABI version mismatch")
    _lib = lib
    return lib


def check(code):
    """Raise DtmmxError for a non-zero return code."""
    if code != DTMMX_OK:
        msg = load().dtmmx_last_error()
        raise DtmmxError(code, msg.decode() if msg else "")
    return code


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def iptr(a):
    """ctypes int* of a C-contiguous int32 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
