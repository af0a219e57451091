"""Python wrapper of libspparks_hip.so with the reference's API
(python/spparks.py:19-95): spparks(name, cmdargs).command / file / extract /
energy / close.  The library runs the strict colour sweep on the GPU."""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")


class SpparksError(RuntimeError):
    pass


class spparks:  # noqa: N801  (name kept from the reference wrapper)
    def __init__(self, name="", cmdargs=None):
        path = os.path.join(LIB_DIR, "libspparks_hip.so" if not name else "libspparks_%s.so" % name)
        if not os.path.exists(path):
            raise SpparksError("Could not load SPPARKS dynamic library %s (build it with "
                               "python -m spparks_amd.build)" % path)
        self.lib = C.CDLL(path, C.RTLD_GLOBAL)
        self.lib.spparks_command.restype = C.c_char_p
        self.lib.spparks_command.argtypes = [C.c_void_p, C.c_char_p]
        self.lib.spparks_file.argtypes = [C.c_void_p, C.c_char_p]
        self.lib.spparks_extract.argtypes = [C.c_void_p, C.c_char_p]
        self.lib.spparks_energy.argtypes = [C.c_void_p]
        self.lib.spparks_energy.restype = C.c_double
        self.lib.spparks_error_message.argtypes = [C.c_void_p]
        self.lib.spparks_error_message.restype = C.c_char_p
        self.lib.spparks_close.argtypes = [C.c_void_p]
        self.spk = C.c_void_p()
        if cmdargs:
            args = ["spparks.py"] + list(cmdargs)
            cargs = (C.c_char_p * len(args))(*[a.encode() for a in args])
            self.lib.spparks_open_no_mpi(len(args), cargs, C.byref(self.spk))
        else:
            self.lib.spparks_open_no_mpi(0, None, C.byref(self.spk))

    def __del__(self):
        if getattr(self, "spk", None):
            self.lib.spparks_close(self.spk)
            self.spk = None

    def close(self):
        self.lib.spparks_close(self.spk)
        self.spk = None

    def _raise_on_error(self):
        msg = self.lib.spparks_error_message(self.spk)
        if msg:
            raise SpparksError(msg.decode())

    def file(self, file):
        self.lib.spparks_file(self.spk, file.encode("utf-8"))
        self._raise_on_error()

    def command(self, cmd):
        r = self.lib.spparks_command(self.spk, cmd.encode("utf-8"))
        if r is None:
            self._raise_on_error()
        return r.decode() if r else ""

    def extract(self, name, type):  # noqa: A002 (argument name from the reference)
        """type as in the reference wrapper: 0 int, 1 int*, 2 int**, 3 double,
        4 double*, 5 double** (e.g. extract("xyz", 5)[i][1])."""
        name = name.encode("utf-8")
        restype = {0: C.POINTER(C.c_int), 1: C.POINTER(C.c_int), 2: C.POINTER(C.POINTER(C.c_int)),
                   3: C.POINTER(C.c_double), 4: C.POINTER(C.c_double),
                   5: C.POINTER(C.POINTER(C.c_double))}.get(type)
        if restype is None:
            return None
        self.lib.spparks_extract.restype = restype
        ptr = self.lib.spparks_extract(self.spk, name)
        if not ptr:
            self._raise_on_error()
            return None
        return ptr[0] if type in (0, 3) else ptr

    def energy(self):
        e = self.lib.spparks_energy(self.spk)
        self._raise_on_error()
        return e
