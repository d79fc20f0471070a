"""Exceptions, mirroring Python/tigre/utilities/errors.py:1-25."""


class Error(Exception):
    """Base class for exceptions in this module."""


class TigreCudaCallError(Error):
    """A native call failed; indexed against the status codes of
    Common/CUDA/errors.hpp:1-7 (the same numbering as include/tigre_amd.h)."""

    error_list = [
        "GPU: HIP error with exit code 1.",
        "There are no available device(s) that support HIP",
        "GPU: One (or more) of your GPUs is being heavily used by another program.",
        "GPU: Assertion Failed.",
        "GPU: Assertion Failed. Memory needed calculation broken!",
        "Invalid argument.",
    ]

    def __init__(self, txt, error_index, detail=""):
        Error.__init__(self)
        self.error_index = error_index
        self.txt = txt
        self.detail = detail

    def __str__(self):
        i = self.error_index - 1
        base = self.error_list[i] if 0 <= i < len(self.error_list) else "error %d" % self.error_index
        return "\n".join(x for x in (self.txt, base, self.detail) if x)
