"""The sparse-set file reader (src/read_files.f90:660-724): record counting,
dim = records // n_sparse, component-fastest layout, bit-exact round trip."""
import numpy as np
import pytest

from turbogap_amd import sparse_io as IO


def test_round_trip_bit_exact(tmp_path):
    rng = np.random.default_rng(3)
    al = rng.normal(scale=1e-3, size=7)
    Qs = rng.normal(size=(7, 11))
    fd, fa = str(tmp_path / "desc"), str(tmp_path / "alphas")
    IO.write_alphas_and_descriptors(fd, fa, al, Qs)
    a2, Q2 = IO.read_alphas_and_descriptors(fd, fa)
    assert np.array_equal(a2, al) and np.array_equal(Q2, Qs)


def test_component_fastest_layout(tmp_path):
    fd, fa = tmp_path / "desc", tmp_path / "alphas"
    fa.write_text("1.0\n2.0\n")
    fd.write_text("\n".join(str(v) for v in [10, 11, 12, 20, 21, 22]) + "\n")
    al, Qs = IO.read_alphas_and_descriptors(str(fd), str(fa))
    assert al.tolist() == [1.0, 2.0]
    assert Qs.tolist() == [[10, 11, 12], [20, 21, 22]]


def test_empty_alphas_rejected(tmp_path):
    (tmp_path / "a").write_text("")
    (tmp_path / "d").write_text("1\n")
    with pytest.raises(ValueError):
        IO.read_alphas_and_descriptors(str(tmp_path / "d"), str(tmp_path / "a"))
