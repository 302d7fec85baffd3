"""Reference-counted wrapper of a C-ABI handle (TACSObject semantics,
src/TACSObject.h:112-140: the wrapper holds one reference)."""
from ._lib import lib, TACSError


class TACSHandle:
    __slots__ = ("ptr", "__weakref__", "_keep")

    def __init__(self, ptr, keep=()):
        if not ptr:
            msg = lib().tacs_last_error()
            raise TACSError(msg.decode() if msg else "object creation failed")
        self.ptr = ptr
        self._keep = keep
        lib().tacs_incref(ptr)

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            try:
                lib().tacs_decref(p)
            except Exception:
                pass
            self.ptr = None

    @property
    def _as_parameter_(self):
        return self.ptr
