"""The hand-written streaming copy (mrb_stream_copy) that measures the
HBM ceiling: copies exactly, for sizes that leave a partial unrolled
iteration and a grid-stride tail, and rejects misaligned sizes."""
import pytest

pytestmark = pytest.mark.gpu


def test_stream_copy_exact(engine):
    import torch

    from marbles_amd import _lib
    from marbles_amd._lib import MrbError, check

    lib = _lib.load()
    sh = torch.cuda.current_stream().cuda_stream
    for nbytes, blocks in ((16, 0), (4096 * 16 + 48, 3), (1 << 22, 0), ((1 << 22) + 16 * 7, 1024)):
        a = torch.randint(0, 255, (nbytes,), dtype=torch.uint8, device="cuda")
        b = torch.zeros_like(a)
        check(lib.mrb_stream_copy(b.data_ptr(), a.data_ptr(), nbytes, blocks, sh))
        torch.cuda.synchronize()
        assert torch.equal(a, b), (nbytes, blocks)
    with pytest.raises(MrbError):
        check(lib.mrb_stream_copy(b.data_ptr(), a.data_ptr(), 24, 0, sh))
