"""ctypes loader of libsmilei_amd.so (the HIP library built from csrc/).

The library is required: there is no CPU fallback.  Loading fails with a
RuntimeError if it has not been built, or if it was built from sources other
than the ones beside it (the build embeds a hash of csrc/ and the header).
"""
import ctypes
import os

from . import srchash

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmilei_amd.so")

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_F = ctypes.POINTER(ctypes.c_float)
_VP = ctypes.c_void_p
_I3 = ctypes.c_int * 3
_D3 = ctypes.c_double * 3


class SmiParams(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_char_p),
        ("interpolation_order", ctypes.c_int),
        ("interpolator", ctypes.c_char_p),
        ("pusher", ctypes.c_char_p),
        ("maxwell_solver", ctypes.c_char_p),
        ("n_cells", ctypes.c_int * 3),
        ("patch_size", ctypes.c_int * 3),
        ("cell_length", ctypes.c_double * 3),
        ("timestep", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("deposit", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


# Every symbol include/smilei_amd.h declares, with its ctypes signature.
SIGNATURES = {
    "smi_last_error": (ctypes.c_char_p, []),
    "smi_build_hash": (ctypes.c_char_p, []),
    "smi_create": (ctypes.c_int, [ctypes.POINTER(SmiParams), ctypes.POINTER(_VP)]),
    "smi_destroy": (ctypes.c_int, [_VP]),
    "smi_add_species": (ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_int)]),
    "smi_set_particles": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int64] + [_D] * 7),
    "smi_init_species": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_uint64]),
    "smi_init_species_mj": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_uint64]),
    "smi_init_species_profile_mj": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_uint64]),
    "smi_init_species_like": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_uint64]),
    "smi_init_species_profile": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_uint64]),
    "smi_num_particles": (ctypes.c_int64, [_VP, ctypes.c_int]),
    "smi_get_particles": (ctypes.c_int, [_VP, ctypes.c_int] + [_D] * 7 + [_I32, _I32]),
    "smi_patch_counts": (ctypes.c_int, [_VP, ctypes.c_int, _I32]),
    "smi_set_field": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "smi_get_field": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "smi_step": (ctypes.c_int, [_VP, ctypes.c_int]),
    "smi_synchronize": (ctypes.c_int, [_VP]),
    "smi_set_time_frozen": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double]),
    "smi_set_step_count": (ctypes.c_int, [_VP, ctypes.c_int64]),
    "smi_step_count": (ctypes.c_int64, [_VP]),
    "smi_energies": (ctypes.c_int, [_VP, _D]),
    "smi_scalars": (ctypes.c_int, [_VP, _D]),
    "smi_rho": (ctypes.c_int, [_VP, _D]),
    "smi_gauss": (ctypes.c_int, [_VP, _D]),
    "smi_profile_step": (ctypes.c_int, [_VP, _F]),
    "smi_stream": (_VP, [_VP]),
    "smi_set_tuning": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "smi_get_tuning": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "smi_autotune": (ctypes.c_int, [_VP, ctypes.c_int, _F]),
    "smi_interpolate_3d": (ctypes.c_int, [ctypes.c_int] + [_VP] * 6 + [_I3, _I3, _D3, ctypes.c_int64] + [_VP] * 3
                           + [_VP] * 4 + [_VP]),
    "smi_interpolate_2d": (ctypes.c_int, [ctypes.c_int] + [_VP] * 6 + [ctypes.c_int * 2, ctypes.c_int * 2,
                                                                        ctypes.c_double * 2, ctypes.c_int64]
                           + [_VP] * 2 + [_VP] * 4 + [_VP]),
    "smi_push_boris": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int64] + [_VP] * 6 + [_VP] * 3 + [_VP]),
    "smi_currents_3d": (ctypes.c_int, [ctypes.c_int] + [_VP] * 3 + [_I3, _I3, _D3, ctypes.c_double, ctypes.c_double,
                                                                    ctypes.c_int64] + [_VP] * 4 + [_VP] * 2
                        + [ctypes.c_int, _VP]),
    "smi_currents_2d": (ctypes.c_int, [ctypes.c_int] + [_VP] * 3 + [ctypes.c_int * 2, ctypes.c_int * 2,
                                                                    ctypes.c_double * 2, ctypes.c_double,
                                                                    ctypes.c_double, ctypes.c_int64]
                        + [_VP] * 5 + [_VP] * 2 + [ctypes.c_int, _VP]),
    "smi_maxwell_ampere_3d": (ctypes.c_int, [_VP] * 9 + [_I3, ctypes.c_double, _D3, _VP]),
    "smi_maxwell_faraday_3d": (ctypes.c_int, [_VP] * 6 + [_I3, ctypes.c_double, _D3, _VP]),
    "smi_density_3d": (ctypes.c_int, [ctypes.c_int, _VP, _I3, _I3, _D3, ctypes.c_double, ctypes.c_int64]
                       + [_VP] * 4 + [_VP]),
    "smi_density_2d": (ctypes.c_int, [ctypes.c_int, _VP, ctypes.c_int * 2, ctypes.c_int * 2, ctypes.c_double * 2,
                                      ctypes.c_double, ctypes.c_int64] + [_VP] * 3 + [_VP]),
    "smi_maxwell_ampere_2d": (ctypes.c_int, [_VP] * 9 + [ctypes.c_int * 2, ctypes.c_double, ctypes.c_double * 2,
                                                         _VP]),
    "smi_maxwell_faraday_2d": (ctypes.c_int, [_VP] * 6 + [ctypes.c_int * 2, ctypes.c_double, ctypes.c_double * 2,
                                                          _VP]),
    "smi_current_deposition_3d": (ctypes.c_int, [_VP] * 3 + [ctypes.c_int] * 3 + [_VP] * 6
                                  + [ctypes.c_uint] * 3 + [_VP] * 3 + [ctypes.c_uint] + [ctypes.c_double] * 7
                                  + [ctypes.c_int] * 6 + [ctypes.c_bool]),
    "smi_density_deposition_3d": (ctypes.c_int, [_VP, ctypes.c_int] + [_VP] * 6 + [ctypes.c_uint] * 3 + [_VP] * 3
                                  + [ctypes.c_uint] + [ctypes.c_double] * 7 + [ctypes.c_int] * 6
                                  + [ctypes.c_bool]),
    "smi_sum_j_ghosts": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _I3, _I3, _VP, _VP]),
    "smi_exchange_b": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _I3, _I3, _VP, _VP]),
    "smi_sort_exchange": (ctypes.c_int, [ctypes.c_int, _I3, _I3, _D3, ctypes.c_int64] + [_VP] * 7 + [_VP] * 4
                          + [_VP]),
    "smi_sort_exchange_noalloc": (ctypes.c_int, [ctypes.c_int, _I3, _I3, _D3, ctypes.c_int64] + [_VP] * 7
                                  + [_VP] * 4 + [_VP]),
    "smi_sort_exchange_noalloc_max_words": (ctypes.c_int64, []),
}

_lib = None
_TAG = b"smilei-amd-source-hash:"


def built_hash(path=LIB_PATH):
    """Source hash embedded in the built library (read from the file, so a
    stale library is found before it is loaded), or None."""
    if not os.path.exists(path):
        return None
    with open(path, "rb") as f:
        data = f.read()
    i = data.find(_TAG)
    if i < 0:
        return None
    return data[i + len(_TAG):i + len(_TAG) + 16].decode("ascii", "replace")


def stale():
    """True if the library is missing or was built from other sources."""
    return built_hash() != srchash.source_hash()


def load():
    """Load libsmilei_amd.so and attach the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libsmilei_amd.so is not built (%s); run __graft_entry__.build() or make -C smilei_amd/csrc" % LIB_PATH)
    have, want = built_hash(), srchash.source_hash()
    if have != want:
        raise RuntimeError("libsmilei_amd.so was built from other sources (library %s, tree %s); rebuild it with "
                           "__graft_entry__.build() or make -C smilei_amd/csrc" % (have, want))
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError("smilei_amd error %d: %s" % (rc, load().smi_last_error().decode()))
    return rc
