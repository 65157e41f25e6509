"""Energy conservation, checked in FP64 on the device and reported in the
stats: every photon of a build ends absorbed, escaped or dropped (the
integer identity photons x weights = absorbed + escaped + dropped holds
exactly, the FP64 figures within the FP32 rounding of the photon weights),
and a
converged solve balances absorbed against radiated, conducted and stored
power.  The figures are compared with numpy sums of the operator's column
sums, the assembled sources and the oracle's photon counts."""
import numpy as np
import pytest

from oracle import physics

pytestmark = pytest.mark.gpu

SIG = 5.670374419e-8


@pytest.mark.parametrize("which,layout", [("box", 0), ("plates", 0), ("city", 0), ("city", 1)])
def test_build_energy_balance(irtt, oracle, which, layout):
    from irtt import scenes
    sd = {"box": lambda: scenes.closed_box(3, emissivity=0.7),
          "plates": lambda: scenes.parallel_plates(n=8, emissivity=0.9),
          "city": lambda: scenes.city_block(seed=2, nx=2, ny=2, target_tris=2000)}[which]()
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        st = s.build_transport(threads_per_tri=64, rays_per_thread=4, batch_seed=3, layout=layout)
        colsum = s.column_sums()
        _, ea, _ = s.bin_data()
    sc = oracle.OracleScene.from_irtt(sd)
    _, _, ost = oracle.transport(sc, oracle.make_params(64, 4, batch_seed=3, bin_mode=0))
    w, q, Q = oracle.photon_weights(sc, oracle.make_params(64, 4, batch_seed=3, bin_mode=0))
    print(f"{which} layout {layout}: emitted {st['emitted_w']:.6e} absorbed {st['absorbed_w']:.6e} "
          f"escaped {st['escaped_w']:.6e} dropped {st['dropped_w']:.6e} rel {st['energy_rel']:.2e}")
    assert st["energy_exact"] == 1
    # the photon weights are FP32 (sigma eps A / photons, as the reference's
    # shader forms them): a few FP32 roundings against the FP64 emitted power
    assert st["energy_rel"] < 3e-7
    # emitted = sum of sigma eps A over the bins
    assert abs(st["emitted_w"] - SIG * ea.sum()) <= 1e-12 * st["emitted_w"]
    # what the operator does not absorb is what escaped or was dropped (up to
    # the same weight rounding)
    assert abs(colsum.sum() - (st["escaped_w"] + st["dropped_w"])) <= 3e-7 * st["emitted_w"]
    if which == "box":
        assert st["escaped_w"] == 0.0 and ost["miss"] == 0
    else:
        assert (st["escaped_w"] > 0.0) == (ost["miss"] > 0)


@pytest.mark.parametrize("steady", [True, False])
def test_solve_power_balance(irtt, steady):
    """A converged solve's power balance over the free bins; power_in equals
    the numpy sum of the assembled sources plus the ambient term."""
    from irtt import scenes
    sd = scenes.parallel_plates(n=8, emissivity=0.9, kelvin=293.0)
    sd.heat_capacity[:] = 2e4
    sd.conductivity[:] = 0.5
    sd.fixed[:10] = 1
    with irtt.Scene.from_data(sd, irtt.BIN_VERTEX) as s:
        s.build_transport(threads_per_tri=128, rays_per_thread=4, batch_seed=1)
        colsum = s.column_sums()
        st = s.solve(irtt.solver_params(steady=steady, dt=300.0, ambient_kelvin=293.0,
                                        x_tol=1e-10, f_tol=1e-20))
    _, _, _, src = physics.bins(sd, 0)
    free = sd.fixed == 0
    want_in = (src + 293.0 ** 4 * colsum)[free].sum()
    print(f"steady={steady}: in {st['power_in']:.6f} W out {st['power_out']:.6f} W "
          f"rel {st['balance_rel']:.2e}; {st}")
    assert st["converged"]
    assert abs(st["power_in"] - want_in) <= 1e-12 * abs(want_in)
    assert st["balance_rel"] < 1e-8
