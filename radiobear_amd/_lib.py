"""ctypes binding of librbx.so (the C ABI declared in include/rbx.h).

The library is loaded from the package directory; there is no fallback: if it
is missing or fails to load, every call raises ``RbxLibraryError``.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'librbx.so')
ABI_VERSION = 10
# rows of rbx_atmosphere_props (RBX_PROP_*: the reference's property-table order)
PROP_ROWS = ('Z', 'R', 'P', 'GM', 'AMU', 'REFR', 'N', 'H', 'LAPSE', 'LAPSEP', 'g')

FP64 = 0
FP32 = 1
PRECISIONS = {'fp64': FP64, 'fp32': FP32}

(LINES_NH3_INV, LINES_NH3_ROT, LINES_NH3_V2, LINES_NH3_SJS, LINES_H2S, LINES_PH3, LINES_H2O, LINES_CO,
 TABLE_H2_ORTON) = range(9)
ORTON_NTEMP = 10
LINE_COLUMNS = {
    LINES_NH3_INV: ('fo', 'Io', 'Eo', 'gammaNH3o'),
    LINES_NH3_ROT: ('fo_rot', 'Io_rot', 'Eo_rot', 'gNH3_rot', 'gH2_rot', 'gHe_rot'),
    LINES_NH3_V2: ('fo_v2', 'Io_v2', 'Eo_v2'),
    LINES_NH3_SJS: ('f0', 'I0', 'E', 'G0'),
    LINES_H2S: ('f0', 'I0', 'E', 'GH2S'),
    LINES_PH3: ('f0', 'I0', 'E', 'WgtI0', 'WgtFGB', 'WgtSB'),
    LINES_H2O: ('f_o', 'I_o', 'E_o', 'w_s', 'x_s', 'w_h2', 'x_h2', 'w_he', 'x_he', 'SR'),
    LINES_CO: ('f0', 'I0', 'E'),
    TABLE_H2_ORTON: ('ftab_ghz', 'logv', 'Ttab'),      # lengths n, 6 n ORTON_NTEMP, ORTON_NTEMP
}
CONSTITUENTS = ('h2', 'h2o', 'h2s', 'nh3', 'ph3', 'co', 'clouds')   # RBX_C_* rows
SPECIES = ('H2', 'HE', 'CH4', 'NH3', 'H2O', 'H2S', 'PH3', 'CO')      # RBX_X_* rows of mix
CLOUDS = ('H2O', 'SOLN', 'NH4SH', 'NH3', 'H2S', 'CH4')               # RBX_CL_* rows of the cloud table
CLOUD_FLAGS = ('ice_p', 'water_p', 'nh4sh_p', 'nh3ice_p', 'h2sice_p', 'ch4')   # clouds_idp other_dict keys
CO_SHAPES = {'voigt': 0, 'vvw': 1, 'diff': 2}                         # any other coshape: 3

LAUNCH_FIELDS = ('alpha_k', 'alpha_series', 'alpha_split', 'far_split', 'int_unroll', 'int_map', 'tan_split',
                 'adj_rays', 'adj_copies')

EXPORTS = ('rbx_abi_version', 'rbx_last_error', 'rbx_device_count', 'rbx_catalog_create',
           'rbx_catalog_set_lines', 'rbx_catalog_lines', 'rbx_catalog_destroy', 'rbx_alpha',
           'rbx_alpha_workspace_size', 'rbx_alpha_tangent', 'rbx_raypaths', 'rbx_integrate', 'rbx_integrate_order', 'rbx_integrate_cols', 'rbx_integrate_factors', 'rbx_integrate_profiles',
           'rbx_jacobian', 'rbx_raypaths_batch', 'rbx_integrate_batch', 'rbx_jacobian_batch',
           'rbx_atmosphere_batch', 'rbx_atmosphere_props', 'rbx_alpha_rescale', 'rbx_alpha_noalloc', 'rbx_integrate_adjoint',
           'rbx_integrate_adjoint_batch', 'rbx_alpha_vjp', 'rbx_alpha_tangent_workspace_size',
           'rbx_integrate_adjoint_workspace_size')
NOALLOC_SCRATCH_BYTES = 64 << 20     # RBX_NOALLOC_SCRATCH_BYTES


class RbxLibraryError(RuntimeError):
    """librbx.so is missing or unusable."""


class RbxError(RuntimeError):
    """A librbx call returned a non-zero status."""

    def __init__(self, status, message):
        super().__init__('{} (status {})'.format(message, status))
        self.status = status


class Absorbers(ctypes.Structure):
    _fields_ = [('h2', ctypes.c_int), ('h2o', ctypes.c_int), ('h2s', ctypes.c_int),
                ('nh3', ctypes.c_int), ('ph3', ctypes.c_int), ('co', ctypes.c_int), ('co_shape', ctypes.c_int),
                ('clouds', ctypes.c_int), ('cloud_types', ctypes.c_int)]


class Launch(ctypes.Structure):
    """rbx_launch: the launch shape of one call (0 = the library's built-in choice)."""
    _fields_ = [(n, ctypes.c_int) for n in LAUNCH_FIELDS] + [('reserved', ctypes.c_int * 7)]


def launch_struct(launch):
    """Launch from None / a {field: value} dict / a Launch (None stays None: all built in)."""
    if launch is None or isinstance(launch, Launch):
        return launch
    bad = set(launch) - set(LAUNCH_FIELDS)
    if bad:
        raise ValueError('unknown launch fields {}'.format(sorted(bad)))
    return Launch(**{k: int(v) for k, v in launch.items()})


def launch_ref(launch):
    """ctypes argument for a `const rbx_launch *` (NULL for None)."""
    ls = launch_struct(launch)
    return None if ls is None else ctypes.byref(ls)


class PlanetDesc(ctypes.Structure):
    _fields_ = [('req', ctypes.c_double), ('rpol', ctypes.c_double),
                ('orientation', ctypes.c_double * 2), ('gtype', ctypes.c_int),
                ('limb', ctypes.c_int)]


_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LP = ctypes.POINTER(Launch)
_SIGS = {
    'rbx_abi_version': (_I, []),
    'rbx_last_error': (_I, [ctypes.c_char_p, ctypes.c_size_t]),
    'rbx_device_count': (_I, [ctypes.POINTER(_I)]),
    'rbx_catalog_create': (_I, [ctypes.POINTER(_VP)]),
    'rbx_catalog_set_lines': (_I, [_VP, _I, _I, ctypes.POINTER(ctypes.POINTER(_D))]),
    'rbx_catalog_lines': (_I, [_VP, _I, ctypes.POINTER(_I)]),
    'rbx_catalog_destroy': (_I, [_VP]),
    'rbx_alpha': (_I, [_VP, ctypes.POINTER(Absorbers), _VP, _VP, _VP, _VP, _I, _VP, _VP, _I, _I,
                       _VP, _I, _VP, _VP, _LP, _VP, ctypes.c_size_t, _VP]),
    'rbx_alpha_noalloc': (_I, [_VP, ctypes.POINTER(Absorbers), _VP, _VP, _VP, _VP, _I, _VP, _VP, _I, _I,
                               _VP, _I, _VP, _VP, _LP, _VP]),
    'rbx_alpha_workspace_size': (_I, [_VP, ctypes.POINTER(Absorbers), _I, _I, _I, _I, _LP,
                                      ctypes.POINTER(ctypes.c_size_t)]),
    'rbx_alpha_tangent': (_I, [_VP, ctypes.POINTER(Absorbers), _VP, _VP, _VP, _I, _VP, _VP, _I, _I,
                               _VP, _I, _VP, _VP, _LP, _VP, ctypes.c_size_t, _VP]),
    'rbx_alpha_tangent_workspace_size': (_I, [_VP, ctypes.POINTER(Absorbers), _I, _I, _I, _I, _LP,
                                              ctypes.POINTER(ctypes.c_size_t)]),
    'rbx_raypaths': (_I, [ctypes.POINTER(PlanetDesc), _VP, _VP, _I, _VP, _I, _I, _VP, _I, _VP, _VP]),
    'rbx_integrate': (_I, [_VP, _I, _VP, _I, _I, _VP, _I, _VP, _I, _I, _I, _D, _VP, _VP, _LP, _VP]),
    'rbx_integrate_order': (_I, [_VP, _I, _I, _I, _VP, _I, _VP, _I, _D, _VP, _VP, _VP]),
    'rbx_integrate_cols': (_I, [_VP, _I, _VP, _VP, _I, _I, _VP, _I, _VP, _I, _I, _I, _D, _VP, _VP, _VP, _LP, _VP]),
    'rbx_integrate_factors': (_I, [_VP, _I, _VP, _I, _VP, _I, _VP, _VP]),
    'rbx_integrate_profiles': (_I, [_VP, _I, _VP, _I, _I, _VP, _I, _VP, _I, _I, _I, _VP, _VP, _VP, _VP]),
    'rbx_jacobian': (_I, [_VP, _I, _VP, _VP, _I, _I, _VP, _I, _VP, _I, _D, _VP, _VP, _VP]),
    'rbx_integrate_adjoint': (_I, [_VP, _I, _VP, _I, _I, _VP, _I, _VP, _I, _VP, _I, _D, _VP, _VP, _VP, _LP,
                                   _VP, ctypes.c_size_t, _VP]),
    'rbx_integrate_adjoint_batch': (_I, [_VP, _I, _VP, _I, _I, _VP, _I, _VP, _I, _VP, _VP, _VP, _I, _I, _D, _VP,
                                         _VP, _VP, _LP, _VP, ctypes.c_size_t, _VP]),
    'rbx_integrate_adjoint_workspace_size': (_I, [_I, _I, _I, _I, _LP, ctypes.POINTER(ctypes.c_size_t)]),
    'rbx_alpha_vjp': (_I, [_VP, _VP, _VP, _I, _I, _VP, _VP]),
    'rbx_alpha_rescale': (_I, [_VP, _I, _I, _I, _VP, _I, _VP, _VP]),
    'rbx_atmosphere_batch': (_I, [_VP, _VP, _VP, _VP, _I, _VP, _I, _D, _D, _D, _VP, _VP, _VP, _VP]),
    'rbx_atmosphere_props': (_I, [_VP, _VP, _VP, _VP, _I, _VP, _I, _D, _D, _D, _VP, _VP]),
    'rbx_raypaths_batch': (_I, [ctypes.POINTER(PlanetDesc), _VP, _VP, _I, _VP, _I, _VP, _I, _VP, _I, _VP, _I,
                                _VP, _VP]),
    'rbx_integrate_batch': (_I, [_VP, _I, _VP, _I, _I, _VP, _I, _VP, _I, _VP, _VP, _VP, _I, _I, _I, _D, _VP,
                                 _VP, _LP, _VP]),
    'rbx_jacobian_batch': (_I, [_VP, _I, _VP, _VP, _I, _I, _VP, _I, _VP, _I, _VP, _VP, _VP, _I, _I, _D, _VP,
                                _I, _VP, _VP]),
}

_lib = None


def lib():
    """The loaded library (loads on first use; raises if unavailable)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RbxLibraryError('librbx.so not built at {} (run python -m radiobear_amd.build)'.format(LIB_PATH))
    # torch ships its own HIP runtime with the same soname (libamdhip64.so.7);
    # loading torch first makes librbx bind to that one runtime, so device
    # pointers and streams from torch are valid in the library.
    import torch  # noqa: F401
    try:
        handle = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise RbxLibraryError('cannot load {}: {}'.format(LIB_PATH, e))
    for name, (res, args) in _SIGS.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    if handle.rbx_abi_version() != ABI_VERSION:
        raise RbxLibraryError('librbx ABI {} != {}'.format(handle.rbx_abi_version(), ABI_VERSION))
    _lib = handle
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().rbx_last_error(buf, len(buf))
    return buf.value.decode(errors='replace')


def check(status):
    if status != 0:
        raise RbxError(status, last_error())
    return status
