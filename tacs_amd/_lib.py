"""Loader for tacs_amd/lib/libtacs_hip.so.

The ctypes signatures are read from the C ABI header include/tacs_hip.h, so
the binding and the header cannot drift apart.  The library is the only
implementation: when it is missing, importing the package raises.
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtacs_hip.so")
HEADER = os.path.join(_HERE, "..", "include", "tacs_hip.h")

_SCALARS = {
    "int": ctypes.c_int,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
    "size_t": ctypes.c_size_t,
    "long": ctypes.c_long,
    "unsigned long long": ctypes.c_ulonglong,
    "void": None,
    "hipStream_t": ctypes.c_void_p,
    "tacs_obj": ctypes.c_void_p,
}


def _ctype(decl):
    decl = decl.replace("const ", "").strip()
    if decl == "char*":
        return ctypes.c_char_p
    if "*" in decl:
        return ctypes.c_void_p
    return _SCALARS[decl]


def parse_header(path=HEADER):
    """Returns {name: (restype, [argtypes])} for every function declared in
    the C ABI header."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"typedef struct .*?\}\s*\w+;", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"([A-Za-z_][\w \*]*?[\s\*])(tacs\w+)\s*\(([^;{]*?)\)\s*;", text):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        ret = re.sub(r"\s+", " ", ret).strip().replace(" *", "*")
        argtypes = []
        args = re.sub(r"\s+", " ", args).strip()
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                t = re.match(r"(.*?)(\w+)$", a).group(1).strip()
                t = t.replace(" *", "*").replace("* ", "*")
                argtypes.append(_ctype(t))
        decls[name] = (_ctype(ret.replace(" *", "*")), argtypes)
    return decls


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python tacs_amd/build.py` "
                "(or __graft_entry__.build()); there is no fallback implementation")
        # PyTorch-ROCm ships its own HIP/ROCr runtime.  Two runtimes in one
        # process compete for the device (the second one finds no GPU), so
        # torch, when installed, is loaded first: the library's libamdhip64
        # dependency then resolves to the runtime already in the process and
        # torch tensors, the library and its null stream share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in parse_header().items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def hip_runtime():
    """ctypes handle through which the HIP runtime the library is bound to is
    reached (dlsym on the library's handle searches its dependencies), e.g.
    hip_runtime().hipEventCreate."""
    return ctypes.CDLL(LIB_PATH) if lib() is not None else None


class TACSError(RuntimeError):
    pass


def check(code, what=""):
    if code != 0:
        msg = lib().tacs_last_error()
        msg = msg.decode() if msg else ""
        raise TACSError(f"{what} failed (code {code}): {msg}")
    return code
