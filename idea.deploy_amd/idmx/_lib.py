"""ctypes bindings of libidmx.so (declared in include/idmx.h).

The library is loaded from this directory (built in-tree by idmx/build.py).
There is no fallback: if the shared library is missing, every entry point
raises, so a run never silently computes on the CPU.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libidmx.so')

c_void_pp = C.POINTER(C.c_void_p)
c_double_p = C.POINTER(C.c_double)


class IdmxError(RuntimeError):
    pass


class ScDesc(C.Structure):
    """Mirror of idmx_sc_desc."""
    _fields_ = [('xi', C.c_int), ('estencil', C.c_int), ('dims', C.c_int * 3),
                ('psi_kind', C.c_int), ('SC_G', C.c_double), ('tau', C.c_double),
                ('collision', C.c_int), ('precision', C.c_int), ('rho0', C.c_double),
                ('mrt_A', C.c_void_p), ('mrt_beta', C.c_void_p), ('mrt_gamma', C.c_void_p),
                ('mrt_diag', C.c_void_p)]


class BatchDomain(C.Structure):
    """Mirror of idmx_batch_domain."""
    _fields_ = [('L0', C.c_int), ('L1', C.c_int), ('SC_G', C.c_double), ('tau', C.c_double),
                ('rho0', C.c_double)]


c_int_p = C.POINTER(C.c_int)

# (name, restype, argtypes)
_SIGS = [
    ('idmx_abi_version', C.c_int, []),
    ('idmx_last_error', C.c_int, [C.c_char_p, C.c_size_t]),
    ('idmx_device_count', C.c_int, [C.POINTER(C.c_int)]),
    ('idmx_ctx_create', C.c_int, [C.c_int, c_void_pp]),
    ('idmx_ctx_destroy', C.c_int, [C.c_void_p]),
    ('idmx_sync', C.c_int, [C.c_void_p]),
    ('idmx_device_name', C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ('idmx_alloc', C.c_int, [C.c_void_p, C.c_size_t, c_void_pp]),
    ('idmx_free', C.c_int, [C.c_void_p, C.c_void_p]),
    ('idmx_h2d', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ('idmx_d2h', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ('idmx_d2d', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ('idmx_memset', C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_size_t]),
    ('idmx_center_of_mass_f64', C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                            c_double_p]),
    ('idmx_sum_f64', C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, c_double_p]),
    ('idmx_max_f64', C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, c_double_p]),
    ('idmx_max_abs_diff_f64', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                        c_double_p]),
    ('idmx_max_abs_diff_f64_noalloc', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                                c_double_p]),
    ('idmx_init_radial', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_int_p,
                                   C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    ('idmx_init_flat', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_int_p,
                                 C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]),
    ('idmx_init_cylinder', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_int_p,
                                     C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]),
    ('idmx_init_single_flat', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, c_int_p,
                                        C.c_double, C.c_double, C.c_int, C.c_int]),
    ('idmx_stream_copy_f64', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                       c_double_p]),
    ('idmx_check_u', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int,
                               c_double_p, c_double_p]),
    ('idmx_sc_create', C.c_int, [C.c_void_p, C.POINTER(ScDesc), c_void_pp]),
    ('idmx_sc_destroy', C.c_int, [C.c_void_p]),
    ('idmx_sc_pop_bytes', C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    ('idmx_sc_set_launch', C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    ('idmx_sc_set_walls', C.c_int, [C.c_void_p, C.c_void_p]),
    ('idmx_sc_set_body_force', C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    ('idmx_sc_get_launch', C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ('idmx_sc_init_pop', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ('idmx_sc_init_pop_coeffs', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          c_double_p]),
    ('idmx_sc_set_psi_program', C.c_int, [C.c_void_p, C.c_char_p, C.c_int, c_double_p,
                                          C.c_int]),
    ('idmx_sc_density_psi', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ('idmx_sc_force_collide_stream', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p]),
    ('idmx_sc_run', C.c_int, [C.c_void_p, c_void_pp, c_void_pp, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_int, C.c_int]),
    ('idmx_sc_run_profiled', C.c_int, [C.c_void_p, c_void_pp, c_void_pp, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                       C.POINTER(C.c_float)]),
    ('idmx_sc_moments', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    ('idmx_sc_density_velocity', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    ('idmx_sc_mass_momentum', C.c_int, [C.c_void_p, C.c_void_p, c_double_p, c_double_p]),
    ('idmx_sc_pop_to_f64', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ('idmx_sc_pop_from_f64', C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ('idmx_sc_last_run_ms', C.c_int, [C.c_void_p, c_double_p]),
    ('idmx_batch_create', C.c_int, [C.c_void_p, C.POINTER(BatchDomain), C.c_int, C.c_int,
                                    C.c_int, C.c_int, c_void_pp]),
    ('idmx_batch_destroy', C.c_int, [C.c_void_p]),
    ('idmx_batch_set_launch', C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    ('idmx_batch_get_launch', C.c_int, [C.c_void_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                          c_int_p]),
    ('idmx_batch_set_active', C.c_int, [C.c_void_p, c_int_p]),
    ('idmx_batch_set_density', C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ('idmx_batch_get_density', C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ('idmx_batch_run', C.c_int, [C.c_void_p, C.c_int]),
    ('idmx_batch_check', C.c_int, [C.c_void_p, C.c_void_p]),
    ('idmx_batch_copy_state', C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    ('idmx_batch_symmetrize', C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ('idmx_batch_force_rest', C.c_int, [C.c_void_p, C.c_int]),
    ('idmx_batch_get_populations', C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    ('idmx_batch_last_run_ms', C.c_int, [C.c_void_p, c_double_p]),
    ('idmx_nn_conv3x3_fwd', C.c_int, [C.c_void_p] * 6 + [C.c_int] * 6),
    ('idmx_nn_conv3x3_bwd_data', C.c_int, [C.c_void_p] * 5 + [C.c_int] * 5),
    ('idmx_nn_conv3x3_bwd_weight', C.c_int, [C.c_void_p] * 5 + [C.c_int] * 5),
    ('idmx_nn_relu_bwd', C.c_int, [C.c_void_p] * 4 + [C.c_int64]),
    ('idmx_nn_pool2', C.c_int, [C.c_void_p] * 3 + [C.c_int] * 5),
    ('idmx_nn_up2', C.c_int, [C.c_void_p] * 3 + [C.c_int] * 7),
    ('idmx_nn_copy_channels', C.c_int, [C.c_void_p] * 3 + [C.c_int64] + [C.c_int] * 5),
    ('idmx_nn_add', C.c_int, [C.c_void_p] * 4 + [C.c_int64]),
    ('idmx_nn_mse', C.c_int, [C.c_void_p] * 4 + [C.c_int64, c_double_p]),
    ('idmx_nn_adam', C.c_int, [C.c_void_p] * 6 + [C.c_int64] + [C.c_float] * 4 + [C.c_int]),
    ('idmx_nn_cast', C.c_int, [C.c_void_p] * 3 + [C.c_int64, C.c_int]),
    ('idmx_nn_resize_to_lattice', C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                            C.c_int, C.c_double]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


def lib():
    """The loaded libidmx.so with argtypes set (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise IdmxError('libidmx.so is not built (%s); run idea.deploy_amd/idmx/build.py'
                            % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    lib().idmx_last_error(buf, 1024)
    return buf.value.decode(errors='replace')


def check(rc, what=''):
    """Raise IdmxError with the library's message on a non-zero status."""
    if rc != 0:
        raise IdmxError('%s failed (%d): %s' % (what, rc, last_error()))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args), name)
