"""Host-side behaviour of the request loop that needs no device: argument
checks and admission refusal of a request larger than the HBM budget."""
import numpy as np
import pytest


def _req_scene():
    from irtt import scenes
    return scenes.city_block(seed=2, nx=2, ny=2, target_tris=1500)


def test_request_over_budget_is_refused(irtt):
    sd = _req_scene()
    with irtt.Server(slots=2, memory_budget=1 << 20) as sv:
        with pytest.raises(irtt.IrttError) as e:
            sv.submit(sd)
        assert e.value.code == -5  # IRTT_ERR_LIMIT


def test_bad_requests_are_rejected(irtt):
    sd = _req_scene()
    with irtt.Server(slots=1, memory_budget=1 << 40) as sv:
        with pytest.raises(irtt.IrttError) as e:
            sv.submit(sd, operator=irtt.OP_ACA)
        assert e.value.code == -1
        with pytest.raises(irtt.IrttError) as e:
            sv.submit(sd, bin_mode=7)
        assert e.value.code == -1
        with pytest.raises(irtt.IrttError):
            sv.poll(12345)


def test_height_parametrization_moves_building_vertices_only():
    from irtt import scenes
    sd = scenes.city_block(seed=3, nx=2, ny=3, target_tris=2000)
    v, p, d, n = scenes.height_parametrization(sd)
    owner = sd.meta["vertex_building"]
    assert n == 6 and (owner[v] >= 0).all() and (owner >= 0).sum() == v.size
    # roof vertices move one-for-one with the height, wall feet stay put
    z = sd.xyz[v, 2]
    h = sd.meta["heights"][owner[v]]
    assert np.allclose(d[z == 0, 2], 0.0)
    roof = np.isclose(z, h)
    assert roof.any() and np.allclose(d[roof, 2], 1.0)
    assert (d[:, :2] == 0).all()
