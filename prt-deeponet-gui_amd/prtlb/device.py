"""Device and stream plumbing (PyTorch-ROCm owns device memory and streams)."""
import torch

from ._lib import PrtError, lib


def device(dev=None):
    """The HIP device to run on; raises when none is visible."""
    if dev is not None:
        dev = torch.device(dev)
        if dev.type != "cuda":
            raise PrtError("prtlb runs on a HIP device only, got %s" % dev)
        return dev
    if not torch.cuda.is_available():
        raise PrtError("prtlb needs a HIP device (MI355X); none is visible")
    lib()
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(dev):
    """hipStream_t of torch's current stream on `dev`, as an int."""
    return torch.cuda.current_stream(dev).cuda_stream


def ptr(t):
    """Device pointer of a tensor (None for None)."""
    return None if t is None else t.data_ptr()
