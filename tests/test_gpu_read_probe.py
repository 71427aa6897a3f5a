"""The device read probe behind bench.py's streaming-read ceiling
(nsparse_testing_read_probe): it reads every 16-byte word of the buffer once,
so its XOR of the buffer's 32-bit words equals numpy's; sizes that are not
whole words are refused on the host before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def probe(nsparse, buf, nbytes, out):
    return nsparse._lib.lib.nsparse_testing_read_probe(ctypes.c_void_p(buf.data_ptr()), ctypes.c_size_t(nbytes),
                                                      ctypes.c_void_p(out.data_ptr()), None)


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, 64 << 20])
def test_probe_reads_every_word(nsparse, nbytes):
    import torch

    words = np.random.default_rng(nbytes).integers(0, 1 << 32, nbytes // 4, dtype=np.uint64).astype(np.uint32)
    buf = torch.from_numpy(words.view(np.int32)).cuda()
    out = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    assert probe(nsparse, buf, nbytes, out) == 0
    torch.cuda.synchronize()
    want = np.bitwise_xor.reduce(words)
    assert np.uint32(out.cpu().numpy().view(np.uint32)[0]) == want


def test_probe_rejects_partial_words(nsparse):
    import torch

    buf = torch.zeros(64, dtype=torch.int32, device="cuda")
    out = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert probe(nsparse, buf, 15, out) == 1  # NSPARSE_ERR_INVALID_ARGUMENT
