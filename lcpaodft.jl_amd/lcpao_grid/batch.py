"""Batched split-BF16x3 or FP32 grid H + density over many structures (C ABI
lcpao_batch_*): one merged launch per operation for all structures
(persistent or one workgroup per item), optionally replayed from a captured
HIP graph."""
import ctypes as C

from . import _lib
from ._lib import GridError


class Batch:
    """Batch(contexts): GridContext objects on one device, each holding the
    whole pair list of its structure and its tables."""

    def __init__(self, contexts):
        self.L = _lib.load()
        self.contexts = list(contexts)  # the contexts must outlive the batch
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        self.h = C.c_void_p()
        rc = self.L.lcpao_batch_create(C.byref(self.h), len(self.contexts), arr)
        if rc:
            raise GridError(rc, "lcpao_batch_create failed")

    def _check(self, rc):
        if rc:
            raise GridError(rc, self.L.lcpao_batch_last_error(self.h).decode())

    @staticmethod
    def _ptrs(ts):
        return (_lib.dp * len(ts))(*[C.cast(t.data_ptr(), _lib.dp) for t in ts])

    def assemble(self, dV, dDM, dH, drho, use_graph=True):
        """H (1 channel) and rho (1 spin) of every structure; torch device
        tensors, one per structure in each list."""
        n = len(self.contexts)
        if not (len(dV) == len(dDM) == len(dH) == len(drho) == n):
            raise ValueError("one V, DM, H and rho per structure")
        self.contexts[0]._order()  # the batch stream does not follow torch's stream
        self._check(self.L.lcpao_batch_assemble(self.h, self._ptrs(dV), self._ptrs(dDM), self._ptrs(dH),
                                                self._ptrs(drho), int(bool(use_graph))))

    def set_precision(self, precision="bf16x3"):
        """"bf16x3" (default) or "fp32": the form of the next assemble calls."""
        self._check(self.L.lcpao_batch_set_precision(self.h, _lib.PRECISIONS[precision]))

    def tune(self, sched=-1):
        """0: persistent work queue, 1: one workgroup per item, 2: as 1 with the density in
        items of at most 8 super-tiles, two workgroups per CU; -1: time them at the next call."""
        self._check(self.L.lcpao_batch_tune(self.h, sched))

    def tuned(self):
        s = C.c_int32()
        ms = (C.c_float * 3)()
        self._check(self.L.lcpao_batch_tuned(self.h, C.byref(s), ms))
        return s.value, [round(float(x), 4) for x in ms]

    def sizes(self):
        n, k, m = C.c_int32(), C.c_int32(), C.c_int64()
        self._check(self.L.lcpao_batch_sizes(self.h, C.byref(n), C.byref(k), C.byref(m)))
        return n.value, k.value, m.value

    def close(self):
        if self.h:
            self.L.lcpao_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
