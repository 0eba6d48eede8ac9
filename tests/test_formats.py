"""Fragment-state formats (CPU): the non-executing pickle decoder against
pickles this test writes itself, refusal of anything but plain data, the
colex <-> lex conversion on the Be 6-31G states the reference keeps as test
data (converted once into tests/golden/be631g_states.npz by
tests/golden/make_be631g_fixture.py), and the .npy loaders."""
import os
import pickle

import numpy as np
import pytest

from qodeapplications_amd import formats as F

GOLD = os.path.join(os.path.dirname(__file__), "golden", "be631g_states.npz")


def test_reader_decodes_lists_and_arrays(tmp_path):
    data = [np.arange(6, dtype=np.float64).reshape(2, 3), [515, 66049, -3, 2 ** 40], (1.5, "x", None, True),
            np.array([1, 2, 3], dtype=np.int64), np.asfortranarray(np.eye(3))]
    for proto in (2, 3, 4, 5):
        p = tmp_path / f"d{proto}.pkl"
        p.write_bytes(pickle.dumps(data, protocol=proto))
        got = F.read_pickled_data(str(p))
        assert np.array_equal(got[0], data[0]) and got[0].dtype == np.float64
        assert got[1] == data[1]
        assert got[2] == data[2]
        assert np.array_equal(got[3], data[3]) and got[3].dtype == np.int64
        assert np.array_equal(got[4], data[4])


class _Evil:
    def __reduce__(self):
        return (os.system, ("true",))


def test_reader_refuses_anything_but_data(tmp_path):
    p = tmp_path / "evil.pkl"
    p.write_bytes(pickle.dumps([1, _Evil()]))
    with pytest.raises(ValueError):
        F.read_pickled_data(str(p))
    p.write_bytes(pickle.dumps({"a": 1}))
    with pytest.raises(ValueError):
        F.read_pickled_data(str(p))


def test_colex_lex_round_trip_on_reference_states():
    g = np.load(GOLD)
    for m in (0, 1):
        for key in ("p1", "p0", "m1"):
            packed = g[f"packed_{m}_{key}"]
            z = g[f"coeffs_colex_{m}_{key}"]
            # the stored order is ascending integers: colexicographic
            assert np.all(np.diff(packed) > 0)
            lex, zl = F.colex_to_lex(packed.tolist(), z, 9, 1)
            assert np.array_equal(lex, g[f"configs_{m}_{key}"])
            assert np.array_equal(zl, g[f"coeffs_{m}_{key}"])
            assert np.array_equal(lex, F.lex_configs(9, lex.shape[1], 1))
            back, zb = F.lex_to_colex(lex, zl)
            assert back == packed.tolist()
            assert np.array_equal(zb, z)
            # orthonormal states, all core orbitals filled
            assert np.abs(zl @ zl.T - np.eye(zl.shape[0])).max() < 1e-13
            assert np.all(lex[:, 0] == 0) and np.all((lex == 9).any(axis=1))


def test_colex_refuses_foreign_configs():
    with pytest.raises(ValueError):
        F.colex_to_lex([0b110, 0b101], np.eye(2), 2, 1)      # core orbital 0 empty in 0b110
    with pytest.raises(ValueError):
        F.colex_to_lex([0b101, 0b101], np.eye(2), 2, 0)      # repeated


def test_npy_loaders(tmp_path):
    cfg = F.lex_configs(3, 2)
    Z = np.linalg.qr(np.random.default_rng(0).standard_normal((cfg.shape[0], 4)))[0]
    np.save(tmp_path / "Z_2e.npy", Z)
    np.save(tmp_path / "configs_2e-stable.npy", cfg)
    st = F.load_npy_states(str(tmp_path / "Z_2e.npy"), str(tmp_path / "configs_2e-stable.npy"), n_states=3)
    assert st.coeffs.shape == (3, cfg.shape[0])
    assert np.array_equal(st.configs, cfg)
    assert np.array_equal(st.coeffs, Z.T[:3])
