"""fmmx_sort_noalloc (keys and stable sort with every temporary in a static
device array) against fmmx_sort, the handle-based entry point it varies,
through the C-ABI: bit-exact keys and permutations on two calls in a row,
the second with fewer points, and the out-of-cube flag."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import old_fmm_bem_amd as F  # noqa: E402


def _noalloc(x, lmax, box, stream=None):
    n = x.shape[0]
    xd = torch.from_numpy(np.ascontiguousarray(x.T)).cuda()
    keys = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    perm = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    bad = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream if stream is None else stream
    rc = F.lib().fmmx_sort_noalloc(0, C.c_void_p(s), n, xd[0].data_ptr(), xd[1].data_ptr(), xd[2].data_ptr(), lmax,
                                   *[float(np.float32(b)) for b in box], keys.data_ptr(), perm.data_ptr(),
                                   bad.data_ptr())
    assert rc == 0, F.lib().fmmx_last_error()
    torch.cuda.synchronize()
    return keys[:n].cpu().numpy(), perm[:n].cpu().numpy(), int(bad.item())


def _with_handle(x, lmax, box):
    f = F.FMM(p=4, lmax=lmax, precision="fp64", box=box)
    k, p = f.sort(x)
    f.close()
    return k, p


@pytest.mark.parametrize("lmax", [2, 3, 5, 8])
def test_noalloc_sort_matches_fmmx_sort(lmax):
    rng = np.random.default_rng(10 + lmax)
    box = (0.0, 0.0, 0.0, 1.0)
    for n in (1_500_000, 37_001):  # two calls in a row, the second smaller
        x = rng.random((n, 3), dtype=np.float32)
        x[: n // 3] = x[0]  # many equal keys: stability decides their order
        k0, p0 = _with_handle(x, lmax, box)
        k1, p1, bad = _noalloc(x, lmax, box)
        print(f"lmax={lmax} n={n}: keys equal {np.array_equal(k0, k1)}, perm equal {np.array_equal(p0, p1)}",
              flush=True)
        assert bad == 0
        assert np.array_equal(k0, k1) and np.array_equal(p0, p1)
        assert np.all(np.diff(k1.astype(np.int64)) >= 0)


def test_noalloc_sort_flags_points_outside_the_cube():
    rng = np.random.default_rng(3)
    x = rng.random((1000, 3), dtype=np.float32)
    _, _, bad = _noalloc(x, 4, (0.0, 0.0, 0.0, 1.0))
    assert bad == 0
    x[17, 1] = 1.5
    _, _, bad = _noalloc(x, 4, (0.0, 0.0, 0.0, 1.0))
    assert bad == 1


def test_noalloc_sort_rejects_oversize_without_launching():
    n = (1 << 21) + 1
    rc = F.lib().fmmx_sort_noalloc(0, None, n, None, None, None, 4, 0.0, 0.0, 0.0, 1.0, None, None, None)
    assert rc == 1 and b"FMMX_SORT_NOALLOC_MAX" in F.lib().fmmx_last_error()
