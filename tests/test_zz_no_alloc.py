"""vk_fft2d_noalloc against vk_fft2d through the C ABI: the same transforms
between caller-owned device buffers with the intermediate in the library's
fixed scratch.  Two calls in a row, the second at a smaller N (the scratch
keeps the first call's contents), each forward and inverse, bit for bit
equal to vk_fft2d; sizes beyond the scratch and host pointers are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))


def test_fft2d_noalloc_equals_fft2d():
    import torch
    assert torch.cuda.is_available()
    from quantumvortices_amd import gpe as G
    from quantumvortices_amd import _native as nat
    L = nat.lib()
    for N, seed in ((1024, 1), (256, 2)):
        print("case fft2d_noalloc N=%d" % N, flush=True)
        g = G.GPE(G.params_from_namelist(N, 1e-3, 2.5))
        a = _rand(N, seed)
        ref_k = g.fft(a, -1)
        ref_x = g.fft(ref_k, +1)
        din = torch.from_numpy(a).to("cuda")
        dk = torch.empty_like(din)
        dx = torch.empty_like(din)
        torch.cuda.synchronize()
        nat.check(L.vk_fft2d_noalloc(g.h, -1, nat.dptr(din), nat.dptr(dk)))
        nat.check(L.vk_fft2d_noalloc(g.h, +1, nat.dptr(dk), nat.dptr(dx)))
        torch.cuda.synchronize()
        k, x = dk.cpu().numpy(), dx.cpu().numpy()
        assert np.array_equal(k, ref_k)
        assert np.array_equal(x, ref_x)
        assert np.linalg.norm(x - a) < 1e-13 * np.linalg.norm(a)
        # host memory is refused
        out = np.empty_like(a)
        assert L.vk_fft2d_noalloc(g.h, -1, nat.ptr(a), nat.ptr(out)) == -1
        g.close()
    # beyond the scratch: refused with a clear error
    g = G.GPE(G.params_from_namelist(2048, 1e-3, 2.5))
    din = torch.zeros((2048, 2048), dtype=torch.complex128, device="cuda")
    dout = torch.empty_like(din)
    torch.cuda.synchronize()
    rc = L.vk_fft2d_noalloc(g.h, -1, nat.dptr(din), nat.dptr(dout))
    assert rc == -4 and b"exceeds" in L.vk_last_error()
    g.close()
