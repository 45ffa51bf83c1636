"""ctypes binding of libhfk.so (the C ABI declared in include/hfk.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when libhfk.so is missing.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhfk.so")

HFK_SCREEN_REFERENCE, HFK_SCREEN_NONE = 0, 1
HFK_BOYS_REFERENCE, HFK_BOYS_ACCURATE = 0, 1
HFK_GUESS_CORE, HFK_GUESS_INTENDED, HFK_GUESS_DENSITY, HFK_GUESS_REFERENCE = 0, 1, 2, 3

# every symbol include/hfk.h declares
EXPORTS = (
    "hfk_default_options", "hfk_create", "hfk_destroy", "hfk_set_options", "hfk_eri_flat",
    "hfk_fock_G", "hfk_fock", "hfk_fock_dev", "hfk_density", "hfk_symm_ortho", "hfk_unitary",
    "hfk_overlap", "hfk_hcore", "hfk_default_scf_options", "hfk_scf", "hfk_gradient",
    "hfk_get_stats", "hfk_get_class_stats", "hfk_last_error", "hfk_eri_flat_noalloc",
)


class HfkBasis(C.Structure):
    _fields_ = [
        ("nbf", C.c_int32), ("nprim_total", C.c_int32), ("natom", C.c_int32),
        ("center", C.POINTER(C.c_double)), ("lxyz", C.POINTER(C.c_int32)),
        ("prim_off", C.POINTER(C.c_int32)), ("coef", C.POINTER(C.c_double)),
        ("expo", C.POINTER(C.c_double)), ("atom_of_bf", C.POINTER(C.c_int32)),
        ("Z", C.POINTER(C.c_double)), ("R", C.POINTER(C.c_double)),
    ]


class HfkOptions(C.Structure):
    _fields_ = [
        ("screen_mode", C.c_int32), ("boys_mode", C.c_int32), ("screen_tau", C.c_double),
        ("schwarz_cut", C.c_double), ("fp32_threshold", C.c_double),
    ]


class HfkScfOptions(C.Structure):
    _fields_ = [
        ("guess", C.c_int32), ("max_steps", C.c_int32), ("diis_size", C.c_int32),
        ("diis_start", C.c_double), ("converge", C.c_double), ("charge", C.c_double),
        ("verbose", C.c_int32), ("pad", C.c_int32), ("guess_density", C.POINTER(C.c_double)),
        ("incremental_tau", C.c_double), ("rebuild_every", C.c_int32), ("pad2", C.c_int32),
        ("energy_converge", C.c_double),
    ]


class HfkScfResult(C.Structure):
    _fields_ = [
        ("energy", C.c_double), ("e_nuc", C.c_double), ("max_error", C.c_double),
        ("steps", C.c_int32), ("occupied", C.c_int32), ("seconds", C.c_double),
        ("fock_seconds", C.c_double), ("delta_energy", C.c_double),
    ]


class HfkStats(C.Structure):
    _fields_ = [
        ("nshell", C.c_int64), ("nshell_pair", C.c_int64), ("nprim_pair", C.c_int64),
        ("quartets_launched", C.c_int64), ("func_quartets", C.c_int64),
        ("last_fock_ms", C.c_double), ("last_eri_ms", C.c_double), ("last_gemm_ms", C.c_double),
        ("last_gemm_flops", C.c_double),
    ]


class HfkClassStat(C.Structure):
    _fields_ = [
        ("la", C.c_int32), ("lb", C.c_int32), ("lc", C.c_int32), ("ld", C.c_int32),
        ("prec", C.c_int32), ("solo", C.c_int32), ("shell_quartets", C.c_int64),
        ("prim_quartets", C.c_int64), ("func_quartets", C.c_int64), ("ms", C.c_double),
        ("tasks", C.c_int64), ("grad_ms", C.c_double),
    ]


_lib = None


def lib():
    """Load libhfk.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libhfk.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int32)
    vp = C.c_void_p
    L.hfk_default_options.argtypes = [C.POINTER(HfkOptions)]
    L.hfk_default_options.restype = None
    L.hfk_create.argtypes = [C.POINTER(HfkBasis), C.POINTER(HfkOptions), C.POINTER(vp)]
    L.hfk_destroy.argtypes = [vp]
    L.hfk_set_options.argtypes = [vp, C.POINTER(HfkOptions)]
    L.hfk_eri_flat.argtypes = [vp, dp, C.c_size_t]
    L.hfk_eri_flat_noalloc.argtypes = [vp, vp, C.c_size_t]
    L.hfk_fock_G.argtypes = [vp, dp, dp]
    L.hfk_fock.argtypes = [vp, dp, dp, dp]
    L.hfk_fock_dev.argtypes = [vp, vp, vp, vp]
    L.hfk_density.argtypes = [vp, dp, C.c_int32, dp]
    L.hfk_symm_ortho.argtypes = [vp, dp, dp]
    L.hfk_unitary.argtypes = [vp, dp, dp, dp]
    L.hfk_overlap.argtypes = [vp, dp]
    L.hfk_hcore.argtypes = [vp, dp]
    L.hfk_default_scf_options.argtypes = [C.POINTER(HfkScfOptions)]
    L.hfk_default_scf_options.restype = None
    L.hfk_scf.argtypes = [vp, C.POINTER(HfkScfOptions), C.POINTER(HfkScfResult), dp, dp, dp, dp]
    L.hfk_gradient.argtypes = [vp, dp, dp, dp]
    L.hfk_get_stats.argtypes = [vp, C.POINTER(HfkStats)]
    L.hfk_get_class_stats.argtypes = [vp, C.POINTER(HfkClassStat), C.c_int32, C.POINTER(C.c_int32)]
    L.hfk_last_error.argtypes = [C.c_char_p, C.c_size_t]
    for name in EXPORTS:
        f = getattr(L, name)
        if name not in ("hfk_default_options", "hfk_default_scf_options"):
            f.restype = C.c_int
    _lib = L
    return L


class HfkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hfk error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        lib().hfk_last_error(buf, 1024)
        raise HfkError(rc, buf.value.decode(errors="replace"))
    return rc
