"""Order 18, the reference's second compiled order (mp = 18,
kernel/precalch.cxx:185-261; include/constants.h:12 sets mp): the FP64 engine
in the EXACT convention on the GPU.

  Laplace G, Gni, Gnj  against the oracle's FMM at mp = 18 on the same tree
                       (the oracle at that order is pinned against the
                       direct sum in test_oracle.py);
  Yukawa               against the oracle's FP64 direct sum, potential and
                       gradient, and more accurate than order 12 on the same
                       points.
"""
import numpy as np
import pytest

from conftest import rel_l2, uniform_cube

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import old_fmm_bem_amd as F  # noqa: E402
from oracle import oracle as O  # noqa: E402


@pytest.mark.parametrize("neq", [0, 1, 2])
def test_laplace_order_18_vs_oracle(neq):
    n = 12000
    x, q = uniform_cube(n, 13)
    nrm = np.random.default_rng(5).standard_normal((n, 3)).astype(np.float32)
    print(f"p=18 neq={neq}", flush=True)
    f = F.FMM(p=18, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x, x, ni_normals=nrm, nj_normals=nrm)
    v, g = f.evaluate(q, neq=neq)
    ov, og = O.fmm(x, x, q, mp=18, lmax=3, neq=neq, conv=O.EXACT, box=(0, 0, 0, 1), ni_normals=nrm,
                   nj_normals=nrm, fp64_out=True)
    ev = rel_l2(v, ov)
    print(f"potential vs oracle {ev:.3e}", flush=True)
    assert ev < 1e-11
    if neq == 0:
        assert rel_l2(g, og) < 1e-11
    f.close()


@pytest.mark.parametrize("lam", [0.5, 4.0])
def test_yukawa_order_18_vs_direct(lam):
    n = 30000
    x, q = uniform_cube(n, 60)
    idx = np.random.default_rng(61).choice(n, 600, replace=False)
    dv, dg = O.direct_yukawa(x[idx], x, q, lam, conv=O.EXACT)
    err = {}
    for p in (12, 18):
        print(f"lam={lam} p={p}", flush=True)
        f = F.FMM(p=p, lmax=4, precision="fp64", convention="exact", box=(0, 0, 0, 1), lam=lam)
        f.set_points(x)
        v, g = f.evaluate(q, neq=F.YUKAWA)
        err[p] = (rel_l2(v[idx], dv), rel_l2(g[idx], dg))
        print(f"potential {err[p][0]:.3e} gradient {err[p][1]:.3e}", flush=True)
        f.close()
    assert err[18][0] < 1e-7 and err[18][1] < 5e-7
    assert err[18][0] < 0.1 * err[12][0]


def test_order_18_needs_fp64_exact():
    for precision, convention in [("mixed", "exact"), ("fp64", "ref")]:
        with pytest.raises(F.FMMError):
            F.FMM(p=18, lmax=3, precision=precision, convention=convention, box=(0, 0, 0, 1))
