"""Leaf3D organ meshes (Mesh%create(meshtype="Leaf3D"), MeshClass.f90:6017-6276), host side.

Pins: the outline's closed-form half-widths at the base, the widest point and the tip; the
base thickening to width/10 (default) or x5 (soybean); unchanged z and Cube3D connectivity;
positive Jacobians at every Gauss point of the shaped elements."""
import numpy as np
import pytest

from plantfem_amd.fem import PF_GLYCINE_MAX, FEMDomain, cube3d, leaf3d

CORNERS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)


def min_detj(nod, el):
    g = 1.0 / np.sqrt(3.0)
    out = np.inf
    for gp in CORNERS * g:
        dN = np.stack([CORNERS[:, k] * 0.125 * np.prod(1 + np.delete(CORNERS, k, 1) * np.delete(gp, k), axis=1)
                       for k in range(3)])
        J = np.einsum("ka,eaj->ekj", dN, nod[el - 1])
        out = min(out, np.linalg.det(J).min())
    return out


def half_width(nod, zval):
    on = np.isclose(nod[:, 2], zval, rtol=0, atol=1e-12)
    return np.abs(nod[on, 0]).max(), nod[on, 1].max() - nod[on, 1].min()


def test_leaf3d_default_outline():
    w, t, L, s = 0.2, 0.01, 1.0, 0.3
    nod, el = leaf3d(8, 2, 20, w, t, L, shaperatio=s)
    n0, e0 = cube3d(8, 2, 20, w, t, L)
    assert np.array_equal(el, e0) and np.array_equal(nod[:, 2], n0[:, 2])
    hw_base, th_base = half_width(nod, 0.0)
    assert np.isclose(hw_base, 0.1 * w / 2, rtol=1e-12)
    assert np.isclose(th_base, w / 10, rtol=1e-12)            # base thickened to width/10
    hw_top, th_top = half_width(nod, L)
    assert np.isclose(hw_top, 0.1 * w / 2, rtol=1e-12) and np.isclose(th_top, t, rtol=1e-12)
    hw_mid, _ = half_width(nod, n0[:, 2][np.argmin(np.abs(n0[:, 2] - s * L))])
    assert np.isclose(hw_mid, w / 2, rtol=1e-9)               # full width at z = shaperatio * length
    assert min_detj(nod, el) > 0


def test_leaf3d_soybean_outline():
    w, t, L = 0.1, 0.01, 0.3
    nod, el = leaf3d(8, 2, 30, w, t, L, species=PF_GLYCINE_MAX, SoyWidthRatio=0.5)
    n0, _ = cube3d(8, 2, 30, w, t, L)
    x0 = n0[:, 0] - w / 2
    base = n0[:, 2] <= L / 25
    assert np.allclose(nod[base, 0], x0[base] / 10, rtol=1e-14, atol=0)
    # thickness x5 at the base, then the curl y -= 0.1 theta (theta - 1), theta = |x| / (w/2)
    theta = np.abs(nod[:, 0] / (w / 2))
    y_uncurled = nod[:, 1] + 0.10 * theta * (theta - 1.0)
    assert np.allclose(y_uncurled[base], 5.0 * (n0[base, 1] - t / 2), rtol=1e-12, atol=1e-15)
    assert np.allclose(y_uncurled[~base], n0[~base, 1] - t / 2, rtol=1e-12, atol=1e-15)
    # widest at z = alpha = 0.5 w (arc top), narrowing to the tip
    hw_alpha, _ = half_width(nod, n0[:, 2][np.argmin(np.abs(n0[:, 2] - 0.5 * w))])
    assert np.isclose(hw_alpha, w / 2, rtol=1e-6)
    assert half_width(nod, L)[0] < 0.05 * w
    assert min_detj(nod, el) > 0
    with pytest.raises(ValueError):
        leaf3d(4, 2, 4, species=PF_GLYCINE_MAX, SoyWidthRatio=1.0)


def test_femdomain_create_leaf3d_and_move():
    d = FEMDomain().create("Leaf3D", x_num=6, y_num=2, z_num=10, x_len=0.1, y_len=0.01, z_len=0.2)
    nod, _ = leaf3d(6, 2, 10, 0.1, 0.01, 0.2)
    assert np.array_equal(d.nodcoord, nod)
    d.move(x=1.0, z=-0.5)
    assert np.allclose(d.nodcoord, nod + [1.0, 0.0, -0.5], rtol=0, atol=1e-15)


def test_circle2d_tiles_the_outer_polygon():
    # Circle2D (MeshClass.f90:6524-6743): quads counter-clockwise, conforming, and together
    # exactly the polygon of the outermost used ring (no gaps, no overlaps)
    from plantfem_amd.fem import circle2d, clean_mesh
    for xnum in (4, 10):
        nod, el = clean_mesh(*circle2d(xnum, xnum))
        P = nod[el - 1]                                      # (ne, 4, 2)
        area = 0.5 * np.sum(P[:, :, 0] * np.roll(P[:, :, 1], -1, axis=1) - np.roll(P[:, :, 0], -1, axis=1) * P[:, :, 1],
                            axis=1)
        assert area.min() > 0
        edges = {}
        for q in el:
            for a in range(4):
                e = tuple(sorted((q[a], q[(a + 1) % 4])))
                edges[e] = edges.get(e, 0) + 1
        assert set(edges.values()) <= {1, 2}
        bnd = [e for e, c in edges.items() if c == 1]
        m = 2 * (xnum // 2 + 1)
        assert len(bnd) == 4 * m                              # the outer ring's 8m/2... 4m-gon edges
        r = np.hypot(nod[:, 0], nod[:, 1])
        R = r.max()
        k = len(bnd)
        assert np.isclose(area.sum(), 0.5 * k * R * R * np.sin(2 * np.pi / k), rtol=1e-12)


def test_cylinder_stem_mesh():
    from plantfem_amd.fem import cylinder3d
    nod, el = cylinder3d(10, 10, 6, 0.08, 0.08, 0.6)
    assert np.isclose(nod[:, 0].max() - nod[:, 0].min(), 0.08, rtol=1e-14)
    assert np.isclose(nod[:, 1].max() - nod[:, 1].min(), 0.08, rtol=1e-14)
    assert np.isclose(nod[:, 2].max(), 0.6, rtol=1e-14) and nod[:, 2].min() == 0.0
    assert el.min() == 1 and el.max() == nod.shape[0]                  # cleaned: every node used
    assert min_detj(nod, el) > 0
    d = FEMDomain().create("Cylinder", x_num=10, y_num=10, z_num=6, x_len=0.08, y_len=0.08, z_len=0.6)
    assert np.array_equal(d.nodcoord, nod) and np.array_equal(d.elemnod, el)


def test_cylinder_refuses_what_it_cannot_build():
    """Circle2D with x_num, y_num <= 1 would index ring nodes that do not exist; and "Cylinder3D"
    is FEMDomain_%create's to_cylinder route (FEMDomainClass.f90:7279-7282), not this mesh."""
    import pytest
    from plantfem_amd.fem import FEMDomain, cylinder3d
    for xy in ((1, 1), (0, 1), (1, 0)):
        with pytest.raises(ValueError, match="x_num or y_num"):
            cylinder3d(xy[0], xy[1], 2)
    assert cylinder3d(2, 1, 2)[1].shape[0] > 0
    with pytest.raises(ValueError, match="Cylinder3D"):
        FEMDomain().create("Cylinder3D", x_num=4, y_num=4, z_num=2)


def test_fortran_cylinder_refuses_small_counts(tmp_path):
    """The Fortran FEMDomain_%create("Cylinder") stops with a message at x_num, y_num <= 1
    instead of writing past its ring arrays (host only)."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "plantfem_amd" / "lib" / "bin" / "organ_meshes"
    out = tmp_path / "m.bin"
    res = subprocess.run([str(exe), str(out), "1", "1"], capture_output=True, text=True, timeout=60)
    assert "Cylinder needs x_num or y_num >= 2" in res.stdout + res.stderr
    assert not out.exists()
    ok = subprocess.run([str(exe), str(out), "2", "1"], capture_output=True, text=True, timeout=60)
    assert ok.returncode == 0 and out.exists()
