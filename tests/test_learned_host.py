"""Host-side pieces of the learned guess: sample points and the P1 pressure
projection (no GPU)."""
import numpy as np
import pytest

from goma_amd import learned, problems
from goma_amd.mesh import HEX27_POS, QUAD9_POS
from oracle import shape as oracle_shape


@pytest.mark.parametrize("dim", [2, 3])
def test_q2_basis_matches_oracle_shape(dim):
    rng = np.random.default_rng(dim)
    for _ in range(5):
        xi = rng.uniform(-1, 1, dim)
        phi = learned.q2_basis(dim, xi)
        ref, _ = oracle_shape(dim, *xi, *([0.0] * (3 - dim)))
        assert np.allclose(phi, ref, atol=1e-14)
    pos = QUAD9_POS if dim == 2 else HEX27_POS
    for k, p in enumerate(pos):
        e = np.zeros(len(pos))
        e[k] = 1
        assert np.allclose(learned.q2_basis(dim, p), e)


def _first_unknown(prob):
    """Goma's node-major numbering: (U1..Ud[, T]) per node, P1 dofs after them
    on the element centroid nodes."""
    m = prob.mesh
    nv = m.dim + int(prob.phys.energy)
    cen = set(m.connect[:, 8 if m.dim == 2 else 20].tolist())
    fu = np.zeros(m.num_nodes + 1, np.int64)
    for n in range(m.num_nodes):
        fu[n + 1] = fu[n] + nv + (m.dim + 1 if n in cen else 0)
    return fu


def test_sample_points_layout():
    prob = problems.natural_convection(n=4, Ra=1e3, carreau=True, jitter=0.15)
    pts = learned.sample_points(prob.mesh)
    m = prob.mesh
    assert pts.shape == (m.num_nodes + 5 * m.num_elems, 2)
    e = pts[m.num_nodes:].reshape(m.num_elems, 5, 2)
    # centroid sample = centroid node; +-1/2 points symmetric about it for a bilinear map
    assert np.allclose(e[:, 0], m.coor.T[m.connect[:, 8]], atol=1e-14)
    assert np.allclose(0.5 * (e[:, 1] + e[:, 2]), e[:, 0], atol=1e-14)
    assert np.allclose(0.5 * (e[:, 3] + e[:, 4]), e[:, 0], atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_pressure_projection_is_exact(dim):
    """P0 = p(centroid), P_d = p(+1/2 e_d) - p(-1/2 e_d) reproduces a linear
    pressure field at every sample point of a jittered mesh."""
    prob = (problems.natural_convection(n=4, Ra=1e3, carreau=True, jitter=0.15) if dim == 2 else
            problems.lid_cavity(n=2, Re=10.0, dim=3, energy=True, Ra=1e3, jitter=0.15))
    m = prob.mesh
    fu = _first_unknown(prob)
    N = int(fu[-1])
    pts = learned.sample_points(m)
    a = np.array([0.3, -1.7, 2.2, 0.9][: dim + 1])
    p = a[0] + pts @ a[1:]
    spe = 1 + 2 * dim
    pe = p[m.num_nodes:].reshape(m.num_elems, spe)
    x = np.zeros(N)
    en = int(prob.phys.energy)
    cen = fu[m.connect[:, 8 if dim == 2 else 20]] + dim + en
    x[cen] = pe[:, 0]
    for d in range(dim):
        x[cen + 1 + d] = pe[:, 1 + 2 * d] - pe[:, 2 + 2 * d]
    t, mask = learned.field_values(prob, fu, x)
    sel = mask[:, dim] > 0
    assert np.max(np.abs(t[sel, dim] - p[m.num_nodes:])) < 1e-12
    assert sel.sum() == m.num_elems * spe
