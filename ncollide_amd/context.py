"""Device context: one GPU, one HIP stream (torch's current stream by default)."""
import ctypes as C

import torch

from . import _lib


class Context:
    """Owns an ncx_ctx bound to a device and a stream.

    The library fails loudly when no GPU or no built libncx.so is present:
    there is no CPU fallback for any query.
    """

    def __init__(self, device=0, stream=None):
        if not torch.cuda.is_available():
            raise _lib.NcxError("ncollide_amd needs a GPU: torch.cuda.is_available() is False")
        self.lib = _lib.load()
        self.device = torch.device("cuda", device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        h = C.c_void_p()
        _lib.check(self.lib.ncx_init(C.byref(h), device, C.c_void_p(stream.cuda_stream)))
        self.h = h

    def set_stream(self, stream):
        self.stream = stream
        _lib.check(self.lib.ncx_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def synchronize(self):
        _lib.check(self.lib.ncx_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.ncx_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default
