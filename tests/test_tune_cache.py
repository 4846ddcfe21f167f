"""The batch tuner's on-disk cache (host logic only): keys name the device
architecture, every solver parameter, the particle count, the torque pass
and the input dtype; the file round-trips and foreign files are refused."""
import json
import types

import pytest

from doublyperiodicstokes_amd import StokesParameters
from doublyperiodicstokes_amd.dpstokes import load_tune_cache, save_tune_cache, tune_key


def _fake_solver(N=100, nz=33, w=6):
    par = StokesParameters(viscosity=1.0, Lx=16.0, Ly=16.0, zmin=0.0, zmax=10.0, w=w, w_d=6, nx=32, ny=32,
                           nz=nz, mode="bottom", precision="mixed")
    # no device here: the key falls back to an unknown architecture
    return types.SimpleNamespace(par=par, N=N, device="cuda:0")


def test_keys_separate_configurations():
    base = tune_key(_fake_solver(), True, "torch.float32")
    assert tune_key(_fake_solver(), True, "torch.float32") == base
    assert tune_key(_fake_solver(N=101), True, "torch.float32") != base
    assert tune_key(_fake_solver(nz=34), True, "torch.float32") != base
    assert tune_key(_fake_solver(w=4), True, "torch.float32") != base
    assert tune_key(_fake_solver(), False, "torch.float32") != base
    assert tune_key(_fake_solver(), True, "torch.float64") != base
    k = json.loads(base)
    assert k["N"] == 100 and k["par"]["nz"] == 33 and k["torques"] is True


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    assert load_tune_cache(path) == {}
    a, b = tune_key(_fake_solver(), True, "f32"), tune_key(_fake_solver(N=7), False, "f32")
    save_tune_cache(path, {a: 16, b: 24})
    assert load_tune_cache(path) == {a: 16, b: 24}
    save_tune_cache(path, {a: 8})
    assert load_tune_cache(path) == {a: 8}
    # the temporary file is renamed over the cache, nothing is left beside it
    assert sorted(p.name for p in tmp_path.iterdir()) == ["tune.json"]


def test_foreign_file_is_refused(tmp_path):
    path = tmp_path / "other.json"
    path.write_text(json.dumps({"something": 1}))
    with pytest.raises(ValueError):
        load_tune_cache(str(path))
