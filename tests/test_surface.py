"""Surface reconstruction (the reference's offline surfacer: Source/re.c,
surface_field.cpp, marching_cubes.cpp).

CPU pins of the oracle: each particle's splat integrates to its volume r_p^3
(the field kernel is normalised, and det G undoes the ellipsoid's stretch,
surface_field.cpp:163-166); a symmetric neighbourhood leaves x_bar on the
particle (eq. 6/10); a flat sheet is flattened by exactly k_r = 4 (eq. 15)
along its normal; the marching-cubes case table (generated, not copied)
equals the product's, and on a sphere and on a random field the surface
is closed and consistently oriented, with the sphere's area and volume.
GPU: the device field, the per-particle ellipsoids and the triangles match
the FP64 oracle."""
import math

import numpy as np
import pytest

from graphics_amd import sphgpu as sg
from oracle import oracle as orc


def _edge_closure(tris):
    """Directed mesh edges by rounded vertex coordinates: every edge must be
    matched by its reverse exactly once (closed, consistently oriented)."""
    v = np.round(tris.reshape(-1, 3) / 1e-7).astype(np.int64).reshape(-1, 3, 3)
    key = lambda p: (int(p[0]), int(p[1]), int(p[2]))
    directed = {}
    for t in v:
        for k in range(3):
            a, b = key(t[k]), key(t[(k + 1) % 3])
            directed[(a, b)] = directed.get((a, b), 0) + 1
    bad = [e for e, c in directed.items() if c != 1 or directed.get((e[1], e[0]), 0) != 1]
    return bad


def test_marching_cubes_table_matches_product_and_edges():
    e1, n1, t1 = sg.marching_cubes_table()
    e2, n2, t2 = orc.marching_cubes_table()
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(n1, n2)
    np.testing.assert_array_equal(t1, t2)
    E = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for c in range(256):
        want = sum(1 << k for k, (a, b) in enumerate(E) if ((c >> a) & 1) != ((c >> b) & 1))
        assert e1[c] == want
        used = set(int(x) for x in t1[c][:3 * n1[c]])
        assert used == {k for k in range(12) if (want >> k) & 1}     # every crossed edge, no other
        assert all(x == -1 for x in t1[c][3 * n1[c]:])
    assert n1[0] == n1[255] == 0


def test_marching_cubes_sphere_closed_oriented_area_volume():
    R, c, cell = 0.37, np.array([0.51, 0.48, 0.53]), 0.02
    n = 52
    g = np.arange(n) * cell
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    f = (R - np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)).astype(np.float32)
    tris = orc.marching_cubes(f, (0.0, 0.0, 0.0), cell, iso=0.0)
    assert len(tris) > 1000
    assert not _edge_closure(tris)
    a, b, cc = tris[:, 0], tris[:, 1], tris[:, 2]
    nrm = np.cross(b - a, cc - a)
    cent = (a + b + cc) / 3 - c
    assert np.all(np.einsum("ij,ij->i", nrm, cent) > 0)             # normals out of the sphere
    area = 0.5 * np.linalg.norm(nrm, axis=1).sum()
    vol = np.einsum("ij,ij->i", a - c, np.cross(b - c, cc - c)).sum() / 6
    assert area == pytest.approx(4 * math.pi * R ** 2, rel=0.01)
    assert vol == pytest.approx(4 / 3 * math.pi * R ** 3, rel=0.01)


def test_marching_cubes_random_field_is_closed():
    rng = np.random.default_rng(4)
    f = rng.uniform(-1, 1, (9, 10, 11)).astype(np.float32)
    f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1] = -1, -1, -1, -1, -1, -1   # closed inside the box
    tris = orc.marching_cubes(f, (0.0, 0.0, 0.0), 0.1, iso=0.0)
    assert len(tris) > 100
    assert not _edge_closure(tris)


def _sheet_and_block():
    s = 0.01
    block = sg.lattice((14, 14, 14), (0.3, 0.3, 0.3), s, 0.0, 0, 0)
    return s, block


def test_splat_integrates_to_particle_volume():
    s = 0.01
    mn, mx = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    for aniso in (False, True):
        x = np.array([[0.5, 0.5, 0.5]], np.float32)                   # isolated: G = 3 k_n I
        f, xb, G, det, nb = orc.surface_field(mn, mx, s, 2 * s, s / 4, x, aniso)
        cell = s / 4
        assert f.sum() * cell ** 3 == pytest.approx(s ** 3, rel=2e-3)
        if aniso:
            assert nb[0] == 0 and det[0] == pytest.approx((0.95 * 3) ** 3, rel=1e-12)
    # every particle of a block, with the anisotropic ellipsoids
    _, block = _sheet_and_block()
    f, xb, G, det, nb = orc.surface_field(mn, mx, s, 2 * s, s / 4, block, True)
    assert f.sum() * (s / 4) ** 3 == pytest.approx(len(block) * s ** 3, rel=5e-3)


def test_symmetric_neighbourhood_and_flat_sheet():
    s = 0.01
    mn, mx = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    _, block = _sheet_and_block()
    f, xb, G, det, nb = orc.surface_field(mn, mx, s, 2 * s, s / 4, block, True)
    centre = np.argmin(np.linalg.norm(block - block.mean(0), axis=1))
    assert nb[centre] > 200
    np.testing.assert_allclose(xb[centre], block[centre], atol=1e-9)
    ev = np.linalg.eigvalsh(G[centre])
    assert ev.max() / ev.min() == pytest.approx(1.0, abs=1e-9)       # isotropic inside a cube lattice
    sheet = sg.lattice((30, 1, 30), (0.35, 0.5, 0.35), s, 0.0, 0, 0)
    f, xb, G, det, nb = orc.surface_field(mn, mx, s, 2 * s, s / 4, sheet, True)
    k = np.argmin(np.linalg.norm(sheet - sheet.mean(0), axis=1))
    w, V = np.linalg.eigh(G[k])
    assert w[2] / w[0] == pytest.approx(4.0, rel=1e-9)               # k_r: the normal is squeezed 4x
    assert abs(abs(V[1, 2]) - 1.0) < 1e-9                             # ... along y, the sheet normal
    assert det[k] == pytest.approx(np.prod(w), rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("aniso", [False, True])
def test_gpu_field_and_mesh_match_oracle(gpu, aniso):
    s = 0.01
    mn, mx = (0.0, 0.0, 0.0), (0.4, 0.4, 0.4)
    rng = np.random.default_rng(8)
    blob = sg.lattice((12, 9, 10), (0.1, 0.1, 0.1), s, 0.3, 21, 0)
    stray = rng.uniform(0.05, 0.35, (40, 3)).astype(np.float32)       # isolated and sparse particles
    x = np.concatenate([blob, stray])
    sf = sg.Surfacer(mn, mx, s, 2 * s, s / 4)
    (sf.aniso if aniso else sf.isotro)(x)
    fg = sf.field()
    fo, xb, G, det, nb = orc.surface_field(mn, mx, s, 2 * s, s / 4, x, aniso)
    assert fg.shape == fo.shape
    assert np.abs(fg - fo).max() / np.abs(fo).max() < 1e-5
    if aniso:
        xg, Gg, dg, ng = sf.particles(len(x))
        np.testing.assert_array_equal(ng, nb)
        np.testing.assert_allclose(xg, xb, rtol=0, atol=1e-6)
        np.testing.assert_allclose(Gg, G, rtol=0, atol=2e-5 * np.abs(G).max())
        np.testing.assert_allclose(dg, det, rtol=2e-5)
    # marching cubes on the same (device) field: identical cases, vertices to FP32
    tg = sf.marching_cubes(0.5)
    to = orc.marching_cubes(fg, sf.grid_min, sf.cell, 0.5)
    assert len(tg) == len(to) > 100
    np.testing.assert_allclose(tg, to, rtol=0, atol=1e-6)
    assert not _edge_closure(tg.astype(np.float64))
    sf.close()


@pytest.mark.gpu
def test_gpu_field_run_to_run_within_atomic_rounding(gpu):
    """The splat adds to the field nodes with FP32 atomics, so the order of
    the additions (and the last bits of a node) can change from run to run;
    two runs on the same particles agree to within the rounding of those
    sums (relative 1e-6 of the field's maximum)."""
    s = 0.01
    mn, mx = (0.0, 0.0, 0.0), (0.4, 0.4, 0.4)
    x = sg.lattice((14, 10, 12), (0.1, 0.1, 0.1), s, 0.3, 22, 0)
    fields = []
    for _ in range(2):
        sf = sg.Surfacer(mn, mx, s, 2 * s, s / 4)
        sf.aniso(x)
        fields.append(sf.field().astype(np.float64))
        sf.close()
    assert np.abs(fields[0] - fields[1]).max() <= 1e-6 * np.abs(fields[0]).max()
