"""The full-device stencil SpMV (ygpu_csr_spmv, ygpu_batch_spmv) through
the C-ABI: exact against scipy's CSR product on the reference's residual and
energy matrices, and member-independent over a batch."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from yelmo_amd import generator, velocity
from yelmo_amd.params import SsaParams
from yelmo_amd.velocity import ResidentBatch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ["residual", "energy"])
def test_csr_spmv_matches_scipy(form):
    print(f"case spmv csr {form}", flush=True)
    p = generator.make_problem(57, 24e3, seed=3, dtype=np.float64)
    rng = np.random.default_rng(0)
    n = p.nx * p.ny
    N = np.asfortranarray(10 ** rng.uniform(9, 11, (p.nx, p.ny)))
    bx = np.asfortranarray(10 ** rng.uniform(2, 4, (p.nx, p.ny)))
    by = np.asfortranarray(10 ** rng.uniform(2, 4, (p.nx, p.ny)))
    if form == "residual":
        ap, ai, av = oracle.assemble(p, bx, by, N)[:3]
    else:
        ap, ai, av = oracle.assemble_energy(p, bx, by, N)[:3]
    A = sp.csr_matrix((av, ai - 1, ap - 1), shape=(2 * n, 2 * n))
    x = rng.standard_normal(2 * n)
    y = velocity.csr_spmv(ap, ai, av, x)
    ref = A @ x
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print("max rel", err, flush=True)
    assert err <= 1e-14


def test_batch_spmv_is_member_independent():
    print("case spmv batch", flush=True)
    par = SsaParams(visc_method=2, beta_method=4, precision="mixed")
    a = generator.make_problem(41, 32e3, seed=1, dtype=np.float64)
    b = generator.make_problem(57, 24e3, seed=2, dtype=np.float64)
    probs = [a.copy(), b.copy(), a.copy()]
    rb = ResidentBatch(probs, par)
    rb.run()
    y64, ms64, by64 = rb.spmv(fp32=False, reps=3, want_y=True)
    y32, ms32, by32 = rb.spmv(fp32=True, reps=3, want_y=True)
    rb.close()
    na, nb = 2 * 41 * 41, 2 * 57 * 57
    assert np.array_equal(y64[:na], y64[na + nb:])
    assert np.array_equal(y32[:na], y32[na + nb:])
    rel = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
    print("fp32 vs fp64", rel, "ms", ms64, ms32, flush=True)
    assert rel <= 1e-5
    assert by64 == 2 * by32 == (41 * 41 * 2 + 57 * 57) * 22 * 8
