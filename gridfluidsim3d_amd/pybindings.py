"""ctypes plumbing for libgfsx.so, in the manner of pyfluid's pybindings
(src/pyfluid/pybindings.py:4-36 and src/pyfluid/pyfluid.py:11-44 of the
reference): every C entry point takes a trailing ``int *err``; a failure
raises ``RuntimeError(<function name> - <message>)``.

The library is loaded from this package's ``lib/`` directory only.  If it is
missing the import fails loudly: there is no CPU fallback.
"""
import ctypes
import os
from ctypes import byref, c_char_p, c_int

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIBNAME = "libgfsx.so"


class LibraryLoadError(RuntimeError):
    pass


_lib = None


def library_path():
    return os.path.join(LIBDIR, LIBNAME)


def load_library():
    """Load (once) and return the HIP library."""
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch ships its own libamdhip64 (same
        # soname, libamdhip64.so.7).  Loading torch first makes this
        # library's dependency resolve to that copy instead of mapping a
        # second runtime, which would leave torch without a device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        path = library_path()
        if not os.path.exists(path):
            raise LibraryLoadError(
                "libgfsx.so is not built (expected at %s); run "
                "__graft_entry__.build() or python -m gridfluidsim3d_amd._build" % path)
        try:
            _lib = ctypes.CDLL(path)
        except OSError as e:
            raise LibraryLoadError("unable to load %s: %s" % (path, e))
    return _lib


def check_success(success, errprefix):
    if not success:
        lib = load_library()
        f = lib.gfsx_get_error_message
        init_lib_func(f, [], c_char_p)
        msg = f()
        raise RuntimeError(errprefix + (msg.decode() if msg else "unknown error"))


def init_lib_func(libfunc, argtypes, restype):
    if libfunc.argtypes is None:
        libfunc.argtypes = argtypes
        libfunc.restype = restype


def execute_lib_func(libfunc, params):
    """Call ``libfunc(*params, &err)`` and raise on failure.  Unlike the
    reference helper, the status is checked for functions with a return
    value as well."""
    err = c_int(0)
    result = libfunc(*params, byref(err))
    check_success(err.value, libfunc.__name__ + " - ")
    return result


def lib_func(name, argtypes, restype=None):
    """Look up ``name`` and declare its arguments (the trailing ``int*``
    included)."""
    f = getattr(load_library(), name)
    init_lib_func(f, list(argtypes) + [ctypes.POINTER(c_int)], restype)
    return f
