"""The oracle's LAPACK-kind inverse (oracle/lapack.c: DGETRF2 + DGETRI with lwork = n): bit for bit
against the reference's own twin of dgetrf_ / dgetri_ (oracle/_ref, compiled from the
reference's sources where they are present), and against what the mathematics fixes."""

import numpy as np
import pytest

from oracle import lapack as L


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def _cases():
    rng = np.random.default_rng(3)
    out = {
        "near_identity": np.eye(12)[None] + 0.1 * rng.standard_normal((40, 12, 12)),  # the reference's own case
        "general_26": rng.standard_normal((60, 26, 26)),  # pivoting on every step
        "general_odd": rng.standard_normal((30, 15, 15)),
        "small": rng.standard_normal((30, 4, 4)),
        "tiny_values": rng.standard_normal((10, 8, 8)) * 1e-300,  # pivots under DLAMCH('S') scale by division
    }
    # convection matrices of real columns (glb3x2 meteo, lmax 26)
    from oracle import convection, physics
    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid("glb3x2"), seed=4)
    sl = (slice(0, 20), slice(30, 40))
    lmax = met.g.lmax
    m = met.m[sl]
    dkg = physics.dkg(met.kvh[sl], m, met.gph[sl], lmax)
    ncol = m.shape[0] * m.shape[1]

    def col(q):
        return q.reshape(ncol, *q.shape[2:])

    out["convection"] = convection.convection_matrix(
        col(m)[:, :lmax], col(met.entu[sl]), col(met.detu[sl]), col(met.entd[sl]), col(met.detd[sl]),
        col(met.cloud_top[sl]), col(met.cloud_lfs[sl]), col(dkg), 1800.0, lmax,
    )  # fmt: skip
    return out


@pytest.fixture(scope="module")
def cases():
    return _cases()


def test_oracle_matches_the_reference_twin(cases):
    if not L.reference_available():
        pytest.skip("oracle/_ref not built (the reference's sources are not here)")
    for name, a in cases.items():
        assert np.array_equal(_bits(L.inverse_rows(a)), _bits(L.reference_inverse_rows(a))), name


def test_the_inverse_inverts(cases):
    for name in ("near_identity", "convection", "general_odd"):
        a = cases[name]
        y = L.inverse_rows(a)
        n = a.shape[1]
        tol = 1e-12 if name != "general_odd" else 1e-9
        assert np.abs(a @ y - np.eye(n)).max() < tol, name


def test_known_answers():
    # a permutation matrix's inverse is its transpose, exactly (every pivot is 1)
    p = np.eye(6)[[3, 0, 5, 1, 4, 2]]
    assert np.array_equal(L.inverse_rows(p[None])[0], p.T)
    # a diagonal matrix: each entry's reciprocal, exactly
    d = np.diag([2.0, -4.0, 0.5, 8.0])
    assert np.array_equal(L.inverse_rows(d[None])[0], np.diag([0.5, -0.25, 2.0, 0.125]))
    # a 2 x 2 with a row swap: [[1, 2], [4, 8 + 1]] -> det 1, inverse [[9, -2], [-4, 1]]
    a = np.array([[1.0, 2.0], [4.0, 9.0]])
    assert np.array_equal(L.inverse_rows(a[None])[0], np.array([[9.0, -2.0], [-4.0, 1.0]]))
    # a singular matrix is refused, as the reference's LapackInverse refuses it
    with pytest.raises(ValueError):
        L.inverse_rows(np.array([[[1.0, 2.0], [2.0, 4.0]]]))


def test_a_dropped_pivot_search_fails_the_twin_check(cases):
    """The check has teeth: Crout without pivoting (the other kind) differs from the LAPACK kind
    on matrices that need pivoting."""
    from oracle import convection

    a = cases["general_26"][:5]
    assert not np.array_equal(_bits(L.inverse_rows(a)), _bits(convection.crout_inverse(a)))
