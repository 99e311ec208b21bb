"""The oracle reproduces the reference's recorded fingerprint digests (CPU)."""

import pytest

from oracle import cases


@pytest.mark.parametrize("name", [case.name for case in cases.CASES])
def test_oracle_reproduces_recorded_digest(name, golden_digests):
    assert cases.run_case(name) == golden_digests["case_sha256"][name]


def test_a_planted_mistake_in_the_oracle_moves_the_digest(golden_digests, monkeypatch):
    """An off-by-one in a single PML coefficient must change the pml_3d digest."""
    from oracle import layer

    original = layer.axis_coefficients

    def skewed(n, faces, offset, dt, dx, order=2, ln_reflection=None, wraps=True):
        kms, sinv, kps = original(n, faces, offset, dt, dx, order, ln_reflection, wraps)
        if offset == 0.5 and faces[0] > 0:
            kms = kms.copy()
            kms[1] = kms[2]
        return kms, sinv, kps

    monkeypatch.setattr(layer, "axis_coefficients", skewed)
    assert cases.run_case("pml_3d") != golden_digests["case_sha256"]["pml_3d"]
