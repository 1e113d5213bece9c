"""GPU parity of the per-solve inputs (ygpu_dyn_inputs) with the oracle's
restatement (oracle/inputs_oracle.inc and oracle_set_ssa_masks), FP32 as the
reference's wp = sp: ice-front codes and SSA face masks equal, slopes and
stresses within one unit in the last place element by element (both sides
evaluate the reference's operations in its order without contracted
multiply-adds).  Generated ice sheets at several sizes and every boundary
family, with partly covered margin cells, plus the hand-built front of the
oracle's pins."""
import numpy as np
import pytest

from oracle import oracle
from tests.test_oracle_pins_inputs import _front_case
from yelmo_amd import generator
from yelmo_amd.params import SsaParams
from yelmo_amd.velocity import calc_dyn_inputs

pytestmark = pytest.mark.gpu

RHO_I, RHO_SW, G = 910.0, 1028.0, 9.81
INT_KEYS = ("mask_frnt", "ssa_mask_acx", "ssa_mask_acy")
FLOAT_KEYS = ("dzsdx", "dzsdy", "taud_acx", "taud_acy", "taul_int_acx", "taul_int_acy")


def _check(fields, dx, dy, bnd, lateral_bc="all", taud_lim=5e5, grad_lim=0.1,
           slope_bg=(0.0, 0.0)):
    par = SsaParams(boundaries=bnd)
    got = calc_dyn_inputs(fields, par, dx, dy, taud_lim, grad_lim, lateral_bc, slope_bg)
    ref = oracle.dyn_inputs(fields["H_ice"], fields["f_ice"], fields["f_grnd"], fields["z_bed"],
                            fields["z_sl"], fields["z_srf"], dx, dy, taud_lim, grad_lim, RHO_I,
                            RHO_SW, G, lateral_bc, slope_bg, bnd)
    for k in INT_KEYS:
        assert np.array_equal(got[k], ref[k]), k
    exact = 0
    for k in FLOAT_KEYS:
        a, b = got[k].astype(np.float64), ref[k].astype(np.float64)
        ulp = np.spacing(np.abs(ref[k])).astype(np.float64)
        assert np.all(np.abs(a - b) <= ulp), (k, np.abs(a - b).max())
        exact += int(np.array_equal(got[k], ref[k]))
    return exact


@pytest.mark.parametrize("n,bnd", [(41, "zeros"), (81, "infinite"), (64, "periodic"),
                                   (57, "MISMIP3D"), (48, "periodic-x"), (50, "mask")])
def test_dyn_inputs_match_oracle(n, bnd):
    print(f"case dyn inputs n={n} {bnd}", flush=True)
    p = generator.make_problem(n, 1536e3 / n, seed=n, dtype=np.float32)
    f = {k: np.asfortranarray(v) for k, v in p.fields.items()}
    # partly covered cells on some margin cells, thinner in proportion
    rng = np.random.default_rng(n)
    margin = (f["f_ice"] == 1.0) & (rng.random(f["f_ice"].shape) < 0.05)
    f["f_ice"] = np.where(margin, np.float32(0.5), f["f_ice"]).astype(np.float32)
    f["H_ice"] = np.where(margin, f["H_ice"] * np.float32(0.5), f["H_ice"]).astype(np.float32)
    exact = _check(f, p.dx, p.dx, bnd)
    print("fields bit-equal:", exact, "of", len(FLOAT_KEYS))


@pytest.mark.parametrize("lateral_bc", ["none", "floating", "marine", "all"])
def test_dyn_inputs_hand_built_front(lateral_bc):
    fi, fg, zb, zsl = _front_case()
    H = np.where(fi > 0, 500.0 * fi, 0.0)
    zs = np.where(zb > -H * RHO_I / RHO_SW, zb + H, H * (1 - RHO_I / RHO_SW))
    f = {k: np.asfortranarray(v, dtype=np.float32) for k, v in
         dict(H_ice=H, f_ice=fi, f_grnd=fg, z_bed=zb, z_sl=zsl, z_srf=zs).items()}
    _check(f, 2e3, 3e3, "zeros", lateral_bc, slope_bg=(1e-3, -2e-3), grad_lim=0.05)


DERIVED_FIELDS = ("taud_acx", "taud_acy", "taul_int_acx", "taul_int_acy")


def _with(p, d):
    q = p.copy()
    for k in DERIVED_FIELDS:
        q.fields[k] = d[k]
    q.state["ssa_mask_acx"] = d["ssa_mask_acx"]
    q.state["ssa_mask_acy"] = d["ssa_mask_acy"]
    return q


def _without(p):
    q = p.copy()
    for k in DERIVED_FIELDS:
        q.fields[k] = None
    q.state["ssa_mask_acx"] = None
    q.state["ssa_mask_acy"] = None
    return q


def _thicken(p, dH):
    """The next solve of a sequence: thickness changed where there is ice,
    with the surface, flotation and grounded fraction that follow."""
    from yelmo_amd import inputs
    q = p.copy()
    f = q.fields
    H = np.where(f["f_ice"] > 0, np.maximum(f["H_ice"] + dH, 0.0), 0.0).astype(np.float32)
    f["H_ice"] = np.asfortranarray(H)
    f["z_srf"] = np.asfortranarray(inputs.calc_z_srf_max(H, f["f_ice"], f["z_bed"], f["z_sl"],
                                                         RHO_I, RHO_SW).astype(np.float32))
    f["H_grnd"] = np.asfortranarray(inputs.calc_H_grnd(H, f["f_ice"], f["z_bed"], f["z_sl"],
                                                       RHO_I, RHO_SW).astype(np.float32))
    f["f_grnd"] = np.asfortranarray(np.where(f["H_grnd"] > 0, 1.0, 0.0).astype(np.float32))
    return q


def test_session_derives_inputs_on_device():
    """A session under -ygpu_inputs device, given the geometry only (driving
    stress, front stress and masks NULL), gives bit for bit the velocity of
    a session given those inputs from ygpu_dyn_inputs, over two solves of a
    sequence whose thickness changes between them."""
    from yelmo_amd.velocity import Session
    print("case session with device inputs 61^2", flush=True)
    p0 = generator.make_problem(61, 24e3, seed=5, dtype=np.float32)
    rng = np.random.default_rng(5)
    dH = np.asfortranarray(10.0 * rng.standard_normal(p0.fields["H_ice"].shape)).astype(np.float32)
    lim = "-ygpu_taud_lim 2e5 -ygpu_grad_lim 0.5 -ygpu_lat_bc all"
    par = SsaParams(precision="mixed")
    par_dev = par.with_(ssa_lis_opt=par.ssa_lis_opt + " -ygpu_inputs device " + lim)
    seq = [p0, _thicken(p0, dH)]
    sh = sd = None
    for k, p in enumerate(seq):
        d = calc_dyn_inputs(p.fields, par, p.dx, p.dy, 2e5, 0.5, "all")
        ph, pd = _with(p, d), _without(p)
        if k == 0:
            sh, sd = Session(ph, par), Session(pd, par_dev)
        ih, idv = sh.solve(ph), sd.solve(pd)
        assert ih.status == 0 and idv.status == 0
        assert ih.ssa_iter_now == idv.ssa_iter_now and ih.ssa_lin_iter == idv.ssa_lin_iter
        for n in ("ux_b", "uy_b", "visc_eff_int", "beta_acx"):
            assert np.array_equal(ph.state[n], pd.state[n]), (k, n)
        print("solve", k, "picard", ih.ssa_iter_now, flush=True)
    sh.close()
    sd.close()


def test_batch_derives_inputs_on_device():
    from yelmo_amd.velocity import ResidentBatch
    print("case batch with device inputs", flush=True)
    probs = [generator.make_problem(n, 1536e3 / n, seed=n, dtype=np.float32) for n in (41, 57)]
    par = SsaParams(precision="mixed")
    par_dev = par.with_(ssa_lis_opt=par.ssa_lis_opt +
                        " -ygpu_inputs device -ygpu_taud_lim 2e5 -ygpu_grad_lim 0.5")
    ph = [_with(p, calc_dyn_inputs(p.fields, par, p.dx, p.dy, 2e5, 0.5, "all")) for p in probs]
    pd = [_without(p) for p in probs]
    for ps, pr in ((ph, par), (pd, par_dev)):
        rb = ResidentBatch(ps, pr)
        rb.run()
        info = rb.download()
        rb.close()
        assert all(i.status == 0 for i in info)
    for a, b in zip(ph, pd):
        assert np.array_equal(a.state["ux_b"], b.state["ux_b"])
        assert np.array_equal(a.state["uy_b"], b.state["uy_b"])
