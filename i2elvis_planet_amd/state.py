"""Host-side state: numpy arrays laid out exactly like the reference globals.

Node fields are FP64 of length N = xnumx*ynumy with node m = m1*ynumy + m2
(headmart.c:87-99); markers are float32 SoA plus an int8 type
(headmart.c:122-127); the boundary tables use the reference's bondm / bondn /
bondv layout (headmart.c:95-97).
"""
import ctypes as C

import numpy as np

from . import abi


class HostState:
    def __init__(self, xnumx, ynumy, marknum, bondnum=1):
        self.xnumx, self.ynumy = int(xnumx), int(ynumy)
        N = self.xnumx * self.ynumy
        self.gx = np.zeros(self.xnumx)
        self.gy = np.zeros(self.ynumy)
        for f in abi.NODE_FIELDS:
            setattr(self, f, np.zeros(N))
        for f in abi.MARKER_FIELDS_F32:
            setattr(self, f, np.zeros(marknum, np.float32))
        self.markt = np.zeros(marknum, np.int8)
        # magnetisation state and time (markmart.c:651-744; in2mart.c sets 0 and -1)
        self.markmg_old = np.zeros(marknum, np.int8)
        self.markmg_time = np.full(marknum, -1.0)
        self.bondm = np.zeros(4 * N, np.int64)
        self.bondn = np.zeros((bondnum, 3), np.int64)
        self.bondv = np.zeros((bondnum, 4))
        self.timestep = 0.0
        self.timesum = 0.0
        self.gxkoef = 1.0

    @property
    def N(self):
        return self.xnumx * self.ynumy

    @property
    def marknum(self):
        return int(self.markx.shape[0])

    def copy(self):
        o = HostState.__new__(HostState)
        for k, v in self.__dict__.items():
            setattr(o, k, v.copy() if isinstance(v, np.ndarray) else v)
        o._keep = None
        return o

    def cstruct(self):
        """ctypes view (pointers into the numpy arrays, which must stay alive)."""
        s = abi.HostStateC()
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        for f in ("gx", "gy") + abi.NODE_FIELDS:
            a = getattr(self, f)
            assert a.dtype == np.float64 and a.flags.c_contiguous
            setattr(s, f, a.ctypes.data_as(dp))
        for f in abi.MARKER_FIELDS_F32:
            a = getattr(self, f)
            assert a.dtype == np.float32 and a.flags.c_contiguous
            setattr(s, f, a.ctypes.data_as(fp))
        assert self.markt.dtype == np.int8
        s.markt = self.markt.ctypes.data_as(C.POINTER(C.c_int8))
        assert self.markmg_old.dtype == np.int8 and self.markmg_time.dtype == np.float64
        s.markmg_old = self.markmg_old.ctypes.data_as(C.POINTER(C.c_int8))
        s.markmg_time = self.markmg_time.ctypes.data_as(dp)
        s.marknum = self.marknum
        self.bondn = np.ascontiguousarray(self.bondn, np.int64)
        self.bondv = np.ascontiguousarray(self.bondv, np.float64)
        s.bondm = self.bondm.ctypes.data_as(C.POINTER(C.c_int64))
        s.bondn = self.bondn.ctypes.data_as(C.POINTER(C.c_int64))
        s.bondv = self.bondv.ctypes.data_as(dp)
        s.bondnum = int(self.bondn.shape[0])
        s.timestep = self.timestep
        s.timesum = self.timesum
        s.gxkoef = self.gxkoef
        return s

    def pull_scalars(self, s):
        self.timestep = s.timestep
        self.timesum = s.timesum
        self.gxkoef = s.gxkoef
